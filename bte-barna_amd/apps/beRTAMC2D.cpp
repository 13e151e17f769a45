// beRTAMC2D front-end on the GPU engine: src/beyondRTA_MC2d.cpp
// (process_input 1299-1462, main 1902-2434) for rectangular single-material
// geometries.
//
// usage: beRTAMC2D <inputfile.xml> nthreads
//
// Input (<beRTAMC2D>): <geometry file/>, <time dt maxtime/>, <material ...>
// with one or more <propagator T file/> children (a .P.eigen.bin over the
// modes after the Gamma modes, or file="synthetic" with density, seed, m and
// precision attributes to build exp(dt B) on the GPU from the synthetic
// operator at T), <Eeff particles|Ed [Tmax Tmin]/>, optional <RTA/>,
// optional <gradient x y/> (else gradients from the box temperatures),
// optional <seed value/>, optional <interpolation rule="bracket"/>.
// Output: one line per step on stdout, "istep time Npart (ed jx jy) per box",
// and properties.msgpack.bin with one frame per step.
#include "front_common.hpp"
#include <chrono>

using namespace front;

struct Saver {
    std::ofstream out;
    std::vector<double> temps;
};

static int frame_cb(void* user, int istep, double time, double Eeff, int nbox, int n, const int32_t* H) {
    auto* s = static_cast<Saver*>(user);
    (void)istep;
    std::string f = pack_frame(time, Eeff, s->temps, nbox, n, H);
    s->out << f.size() << "#" << f;
    return 0;
}

static int run(int argc, char** argv) {
    if (argc != 3) {
        std::cerr << "Usage: " << argv[0] << " <inputfile.xml> nthreads" << std::endl;
        return 1;
    }
    if (std::atoi(argv[2]) < 2) {
        std::cerr << "nthreads must be at least 2" << std::endl;
        return 1;
    }
    auto root = xml::parse_file(argv[1]);
    if (root->name != "beRTAMC2D") throw std::runtime_error(std::string(argv[1]) + ": root is not <beRTAMC2D>");
    const xml::Node* geo = root->child("geometry");
    const xml::Node* tim = root->child("time");
    const xml::Node* mat = root->child("material");
    const xml::Node* eef = root->child("Eeff");
    if (!geo || !tim || !mat || !eef) throw std::runtime_error("<beRTAMC2D> needs geometry, time, material and Eeff");
    if (root->children("material").size() > 1) throw std::runtime_error("This is currently not supported (several materials)");
    if (root->child("spectral")) std::cout << "# <spectral> is read by dist_reader, not by the solver" << std::endl;
    Boxes B = read_geometry(geo->str("file"));
    Material M = read_material(*mat);
    const bool rta = root->child("RTA") != nullptr;
    bte_ctx* ctx = nullptr;
    check(bte_ctx_create(0, &ctx), "GPU context");
    bte_modes* modes = nullptr;
    check(bte_modes_create(ctx, M.n, M.nq, M.omega.data(), M.vx.data(), M.vy.data(), M.w.data(), M.V, M.lattz,
                           &modes),
          "mode table");
    if (!M.sigma.empty()) check(bte_modes_set_sigma(modes, M.sigma.data()), "per-mode sigma");
    bte_geom* geom = make_geom(ctx, B);
    const double dt = tim->num("dt");
    bte_prop* prop = nullptr;
    if (!rta) {
        // every <propagator T file/> child (beyondRTA_MC2d.cpp:1355-1376)
        const auto pns = mat->children("propagator");
        if (pns.empty()) throw std::runtime_error("material needs a <propagator T file/> unless <RTA/> is given");
        std::vector<double> Ps, Ts;
        int64_t rows = -1;
        for (const xml::Node* pn : pns) {
            const double T = pn->num("T");
            const std::string file = pn->str("file");
            std::vector<double> P;
            int64_t r = 0;
            if (file == "synthetic") {
                // synthetic energy-conserving operator at T, P = exp(dt B) on the GPU
                std::vector<double> C(M.n), Bm((size_t)M.n * M.n);
                for (int i = 0; i < M.n; i++) C[i] = kB * be_kernel(M.omega[i], T);
                check(bte_synthetic_B(M.n, C.data(), M.w.data(), pn->num("density", 0.05),
                                      (uint64_t)pn->num("seed", 1), 200, Bm.data()),
                      "synthetic collision operator");
                int m = (int)pn->num("m", 0);
                if (m <= 0) check(bte_krylov_size(M.n, Bm.data(), dt, 1e-8, &m), "Krylov size");
                if (m > 64) throw std::runtime_error("Krylov size " + std::to_string(m) + " above 64; give m");
                const std::string pr = pn->str("precision", "fp64");
                const int prec = pr == "fp32" ? BTE_PREC_FP32 : (pr == "bf16" ? BTE_PREC_BF16 : BTE_PREC_FP64);
                P.resize((size_t)M.n * M.n);
                bte_krylov_report kr{};
                check(bte_krylov_build(ctx, M.n, Bm.data(), dt, m, prec, BTE_ARNOLDI_TEXTBOOK, P.data(), &kr),
                      "Krylov propagator");
                r = M.n;
                std::cout << "# propagator at " << T << " K: synthetic operator, m = " << m << ", " << pr << ", "
                          << kr.build_ms << " ms on the GPU" << std::endl;
            } else {
                P = load_eigen_bin(file, r);
            }
            if (r > M.n) throw std::runtime_error("propagator larger than the mode table");
            if (rows >= 0 && r != rows) throw std::runtime_error("propagators of different sizes");
            rows = r;
            Ps.insert(Ps.end(), P.begin(), P.end());
            Ts.push_back(T);
        }
        check(bte_prop_create_multi(ctx, (int)rows, (int)Ts.size(), Ps.data(), Ts.data(), M.n - (int)rows, &prop),
              "propagator");
    }
    // <init_distribution box file> (beyondRTA_MC2d.cpp:1411-1417): column 0
    // of a tab-separated file, lines holding '#' skipped (get_column,
    // beyondRTA_MC2d.cpp:1270-1291)
    std::vector<int32_t> init_boxes;
    std::vector<double> init_dist;
    int init_len = -1;
    for (const xml::Node* d : root->children("init_distribution")) {
        const int ib = (int)d->num("box");
        const std::string fn = d->str("file");
        std::ifstream f(fn);
        if (!f) throw std::runtime_error("cannot open init_distribution file " + fn);
        std::vector<double> col;
        std::string line;
        while (std::getline(f, line)) {
            if (line.find('#') != std::string::npos) continue;
            std::istringstream ls(line);
            double v;
            if (ls >> v) col.push_back(v);
        }
        if (init_len >= 0 && (int)col.size() != init_len)
            throw std::runtime_error("init_distribution files of different lengths");
        init_len = (int)col.size();
        init_boxes.push_back(ib);
        init_dist.insert(init_dist.end(), col.begin(), col.end());
    }
    if (!init_boxes.empty())
        std::cout << "# init_distribution: " << init_boxes.size()
                  << " given distributions, kept together with the particles of boxes at Tinit != Teq (the "
                     "reference's second init_particles discards them, beyondRTA_MC2d.cpp:2096-2105)"
                  << std::endl;
    bte_bert_cfg cfg{};
    cfg.thickness = M.thickness;
    cfg.dt = dt;
    cfg.ninit = (int)init_boxes.size();
    cfg.init_len = init_len < 0 ? 0 : init_len;
    cfg.init_boxes = init_boxes.empty() ? nullptr : init_boxes.data();
    cfg.init_dist = init_dist.empty() ? nullptr : init_dist.data();
    cfg.maxtime = tim->num("maxtime");
    cfg.Ed = eef->num("Ed", -1.0);
    cfg.Efactor = (uint64_t)eef->num("particles", 0);
    cfg.Tmax = eef->num("Tmax", 0.0);
    cfg.Tmin = eef->num("Tmin", 0.0);
    cfg.rta = rta ? 1 : 0;
    cfg.pcancel_each = 4;
    cfg.seed = root->child("seed") ? (uint64_t)root->child("seed")->num("value") : 1;
    // <gradient x y>: the same gradient in every regular box instead of
    // calculate_gradientT (beyondRTA_MC2d.cpp:1328-1333, 1441-1460)
    if (const xml::Node* gn = root->child("gradient")) {
        cfg.fixed_gradient = 1;
        cfg.gradient[0] = gn->num("x");
        cfg.gradient[1] = gn->num("y");
    }
    // <interpolation rule="bracket"/>: interpolate untabulated box
    // temperatures between the adjacent propagators (engine option)
    if (const xml::Node* in = root->child("interpolation"))
        cfg.prop_interp = in->str("rule", "reference") == "bracket" ? 1 : 0;
    bte_bert* run = nullptr;
    check(bte_bert_create(ctx, modes, geom, prop, &cfg, &run), "beRTAMC2D setup");
    // steps while time <= maxtime (beyondRTA_MC2d.cpp:2138)
    int nst = 0;
    for (double t = 0; t <= cfg.maxtime; t += dt) nst++;
    std::vector<double> props((size_t)(nst + 1) * B.n() * 3);
    std::vector<int64_t> npart(nst + 1);
    Saver sv;
    sv.out.open("properties.msgpack.bin", std::ios::binary);
    if (!sv.out) throw std::runtime_error("cannot write properties.msgpack.bin");
    sv.temps = B.Teq;
    bte_bert_report rep{};
    auto t0 = std::chrono::steady_clock::now();
    check(bte_bert_execute(run, props.data(), npart.data(), nullptr, frame_cb, &sv, &rep), "beRTAMC2D run");
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    sv.out << "0#";
    sv.out.close();
    std::cout << "#Init particles: " << npart[0] << std::endl;
    std::cout << "#Eeff is: " << rep.Eeff << " J" << std::endl;
    double time = 0;
    for (int s = 0; s <= rep.nsteps; s++) {
        std::cout << s << '\t' << time << '\t' << npart[s] << '\t';
        for (int b = 0; b < B.n(); b++) {
            const double* q = &props[((size_t)s * B.n() + b) * 3];
            std::cout << q[0] << '\t' << q[1] << '\t' << q[2] << '\t';
        }
        std::cout << std::endl;
        time += dt;
    }
    std::cout << "# " << rep.nsteps << " steps, " << rep.particle_steps << " particle-steps in " << el << " s"
              << std::endl;
    bte_bert_destroy(run);
    if (prop) bte_prop_destroy(prop);
    bte_geom_destroy(geom);
    bte_modes_destroy(modes);
    bte_ctx_destroy(ctx);
    return 0;
}

int main(int argc, char** argv) {
    try {
        return run(argc, argv);
    } catch (const std::exception& e) {
        std::cerr << "beRTAMC2D: " << e.what() << std::endl;
        return 1;
    }
}
