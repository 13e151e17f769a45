// RTAMC2D front-end on the GPU engine: the transient branch of
// src/RTA_MC2D.cpp (process_input 1806-2055, main 2085-2471) for
// rectangular single-material geometries.
//
// usage: RTAMC2D <inputfile.xml> <nthreads> [nruns]
// (nthreads is accepted for compatibility; the particle loop runs on the GPU)
//
// Input (<RTA_MC2d>): <geometry file/>, <time dt maxtime/>, <material .../>
// (see front_common.hpp for the mode-table attributes), <particles N
// [seed]/>, optional <ballistic/>, <gradient x y/>, <convergence energy flux/>,
// <spectral><resolution ticks/><location bin/>...</spectral>.
// Output per run: temperature_<T0>K_run_<i>.csv, jx_<T0>K_run_<i>.csv,
// jy_<T0>K_run_<i>.csv with rows t = (i + 0.5) dt; with <spectral>, per
// location deltaT_omega_, jx_omega_, jy_omega_ and fd_q_<loc>_<T0>K_run_<i>.csv
// (RTA_MC2D.cpp:2473-2541); with a nonzero gradient the steady branch writes
// steady_<T0>K_run_<i>.csv (rows T, jx, jy).
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
    // <spectral> (RTA_MC2D.cpp:1930-1945): ticks and locations
    int nomega = 0;
    std::vector<int32_t> locs;
    if (const xml::Node* sp = root->child("spectral")) {
        if (const xml::Node* rs = sp->child("resolution")) nomega = (int)rs->num("ticks");
        for (const xml::Node* l : sp->children("location")) locs.push_back((int32_t)l->num("bin"));
        if (!locs.empty() && nomega < 1) throw std::runtime_error("<spectral> needs <resolution ticks/>");
        if (steady && !locs.empty())
            std::cout << "# <spectral> is produced by the transient branch only" << std::endl;
    }
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
    if (!M.sigma.empty()) check(bte_modes_set_sigma(modes, M.sigma.data()), "per-mode sigma");
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
    // rows: (i + 0.5) dt for i < ntime - 1 with ntime = floor(maxtime/dt) + 1
    const int nt = (int)std::floor(cfg.maxtime / cfg.dt + 1.0e-6);
    // temperature_, jx_ and jy_<T0>K_run_<irun>.csv from [rows][nbox] blocks
    auto write_run = [&](int irun, int rows, const double* T, const double* jx, const double* jy) {
        char tag[64];
        std::snprintf(tag, sizeof tag, "%gK_run_%i.csv", cfg.T0, irun);
        const char* names[3] = {"temperature_", "jx_", "jy_"};
        const double* data[3] = {T, jx, jy};
        for (int q = 0; q < 3; q++) {
            std::vector<double> out((size_t)rows * (B.n() + 1));
            for (int i = 0; i < rows; i++) {
                out[(size_t)i * (B.n() + 1)] = (i + 0.5) * cfg.dt;
                for (int b = 0; b < B.n(); b++) out[(size_t)i * (B.n() + 1) + 1 + b] = data[q][(size_t)i * B.n() + b];
            }
            std::cout << "# Writing " << names[q] << tag << std::endl;
            write_csv(std::string(names[q]) + tag, rows, B.n() + 1, out.data());
        }
    };
    // The reference runs its nruns one after another (RTA_MC2D.cpp:2188-2471);
    // here they go to the GPU as one batch: run i keeps the seed seed + i and
    // the random streams of its own single run (keyed by run seed and particle
    // index), so the histories are those of nruns single runs and only the
    // order of the FP64 tally sums differs. Each run's time bins are summed in
    // FP64 LDS per run (tally mode 3) where one bin of its boxes fits, else by
    // device atomics. Spectral locations need single-run handles.
    const bool batched = !steady && nruns > 1 && locs.empty();
    if (batched) {
        std::vector<bte_rta_run_desc> d(nruns);
        for (int irun = 0; irun < nruns; irun++) {
            d[irun] = bte_rta_run_desc{};
            d[irun].geom = geom;
            d[irun].T0 = cfg.T0;
            d[irun].nparticles = cfg.nparticles;
            d[irun].seed = seed0 + (uint64_t)irun;
            d[irun].ballistic = cfg.ballistic;
            d[irun].wscale = 1.0;
        }
        auto t0 = std::chrono::steady_clock::now();
        bte_rta_batch* bt = nullptr;
        check(bte_rta_batch_create(ctx, modes, nruns, d.data(), cfg.thickness, cfg.dt, cfg.maxtime, cfg.sigma_bounce,
                                   0, &bt),
              "RTAMC2D setup");
        if (bte_rta_batch_set_tally(bt, 3) != BTE_OK) check(bte_rta_batch_set_tally(bt, 0), "RTAMC2D tallies");
        const size_t per = (size_t)nt * B.n();
        std::vector<double> T(per * nruns), jx(T.size()), jy(T.size()), Eeff(nruns);
        bte_rta_batch_report rep{};
        check(bte_rta_batch_execute(bt, T.data(), jx.data(), jy.data(), nullptr, Eeff.data(), &rep), "RTAMC2D runs");
        double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        check(bte_rta_batch_destroy(bt), "RTAMC2D teardown");
        if (rep.ntime - 1 != nt) throw std::runtime_error("time bins of the batch differ from the input's");
        std::cout << "# runs 0-" << nruns - 1 << " as one batch: " << rep.events << " events, " << rep.crossings
                  << " crossings, " << el << " s" << std::endl;
        for (int irun = 0; irun < nruns; irun++) {
            std::cout << "# run " << irun << ": Eeff " << Eeff[irun] << std::endl;
            write_run(irun, nt, T.data() + per * irun, jx.data() + per * irun, jy.data() + per * irun);
        }
    }
    for (int irun = 0; irun < nruns && !steady && !batched; irun++) {
        cfg.seed = seed0 + (uint64_t)irun;
        std::vector<double> T((size_t)nt * B.n()), jx(T.size()), jy(T.size());
        bte_rta_report rep{};
        auto t0 = std::chrono::steady_clock::now();
        bte_rta* rr = nullptr;
        check(bte_rta_create(ctx, modes, geom, &cfg, &rr), "RTAMC2D setup");
        if (!locs.empty())
            check(bte_rta_spectral(rr, (int)locs.size(), locs.data(), nomega, nullptr), "spectral decomposition");
        check(bte_rta_execute(rr, T.data(), jx.data(), jy.data(), &rep), "RTAMC2D run");
        double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        const int rows = rep.ntime - 1;
        std::cout << "# run " << irun << ": " << rep.events << " events, " << rep.crossings << " crossings, Eeff "
                  << rep.Eeff << ", " << el << " s" << std::endl;
        write_run(irun, rows, T.data(), jx.data(), jy.data());
        if (!locs.empty()) {
            // RTA_MC2D.cpp:2473-2541: deltaT_omega, jx_omega, jy_omega as
            // (ntime) x (nomega + 1) with the tick row first (time -1) and
            // data row i at time (i + 0.5) dt; fd_q without a time column
            // under a header of the q coordinates (the mode table has none:
            // zeros)
            std::cout << "# Writing spectral decompositions" << std::endl;
            std::vector<double> om(nomega), dT((size_t)rows * nomega), sx(dT.size()), sy(dT.size()),
                fd((size_t)rows * M.nq);
            for (size_t k = 0; k < locs.size(); k++) {
                check(bte_rta_spectral_get(rr, (int)k, om.data(), dT.data(), sx.data(), sy.data(), fd.data()),
                      "spectral decomposition");
                const char* sn[3] = {"deltaT_omega_", "jx_omega_", "jy_omega_"};
                const std::vector<double>* sd[3] = {&dT, &sx, &sy};
                char stag[96];
                std::snprintf(stag, sizeof stag, "%i_%gK_run_%i.csv", (int)locs[k], cfg.T0, irun);
                for (int q = 0; q < 3; q++) {
                    std::vector<double> out((size_t)(rows + 1) * (nomega + 1));
                    out[0] = -1.0;
                    for (int j = 0; j < nomega; j++) out[1 + j] = om[j];
                    for (int i = 1; i <= rows; i++) {
                        double* r = out.data() + (size_t)i * (nomega + 1);
                        r[0] = (i + 0.5) * cfg.dt;
                        for (int j = 0; j < nomega; j++) r[1 + j] = (*sd[q])[(size_t)(i - 1) * nomega + j];
                    }
                    write_csv(std::string(sn[q]) + stag, rows + 1, nomega + 1, out.data());
                }
                std::string head = "# ";
                for (int q = 0; q < M.nq; q++) head += std::to_string(0.0) + (q + 1 < M.nq ? "," : "");
                head += "\n# ";
                for (int q = 0; q < M.nq; q++) head += std::to_string(0.0) + (q + 1 < M.nq ? "," : "");
                head += "\n# ";
                write_csv(std::string("fd_q_") + stag, rows, M.nq, fd.data(), head);
            }
        }
        check(bte_rta_destroy(rr), "RTAMC2D teardown");
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
