// rpcf_run -- the simulation driver: the reference's run_simulation
// (source/main_loop.cpp:15-115) restated over the rpcf C ABI (include/rpcf.h).
//
//   rpcf_run [--lib PATH] [PARAMETER_FILE]
//
// PARAMETER_FILE (default ../input/parameter.txt, as the reference) holds the reference's
// six whitespace-separated lines (RPCF_Paras::read_para, source/operation.cpp:115-138):
//   mx my mz / n_pi_x n_pi_y / Re Ro dt /
//   start_type start_step end_step save_internal save_recovery_internal /
//   output_prefix / recovery_prefix
// and optionally two extension lines (absent: reference behaviour):
//   c0 c1 c2 flow_control precision    base flow U = c0 + c1 y + c2 y^2, flow-rate control, 0 FP64 / 1 mixed
//   seed amplitude kmax degmax kexp [peak_bound]
//                                      seeded initial field (rpcf_seed_state; peak_bound
//                                      0, the default: coefficients x amplitude); without
//                                      this line start_type 0 takes the reference's
//                                      analytic field (init_flow_kernel)
//
// --lib selects the library implementing the ABI (default: librpcf_hip.so beside this
// executable); the same driver runs the CPU oracle for testing.
//
// Outputs at step 0 and every save_internal steps: the reference's velocity and vorticity
// files (<prefix>_velocity.<step>.dat, <prefix>_vortricity.<step>.dat, physical fields on
// the padded grid), the max |field| lines of its statistics(), and a spectral state
// snapshot <prefix>_state.<step>.bin (an addition).  Output prefix "-": no files.
//
// Differences from the reference, by design: errors are reported and end the program
// with a non-zero status instead of exit()/system("pause"); the recovery file is portable
// (sizes in the header, pencil stride nz) and lives at <recovery_prefix>.rcv.
#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/rpcf.h"

namespace {

struct Api {
    void* h = nullptr;
    int (*sizes)(const rpcf_params*, int*, int*, int*, int*, int*);
    int (*create)(const rpcf_params*, rpcf_ctx**);
    void (*destroy)(rpcf_ctx*);
    const char* (*last_error)(const rpcf_ctx*);
    int (*set_state)(rpcf_ctx*, const double*, const double*, long);
    int (*get_state)(rpcf_ctx*, double*, double*, long);
    int (*get_nonlinear)(rpcf_ctx*, int, double*, double*, long);
    int (*set_nonlinear)(rpcf_ctx*, int, const double*, const double*, long);
    int (*step)(rpcf_ctx*, int);
    int (*bootstrap)(rpcf_ctx*);
    int (*seed_state)(const rpcf_params*, uint64_t, double, int, int, double, int, double*, double*, long);
    int (*get_field)(rpcf_ctx*, int, double*, long);
    int (*phys_to_spec)(rpcf_ctx*, const double*, double*, long, int);
    int (*spec_to_phys)(rpcf_ctx*, const double*, double*, long, int);
};

template <typename F>
bool sym(void* h, const char* name, F* out) {
    *out = reinterpret_cast<F>(dlsym(h, name));
    if (!*out) std::cerr << "rpcf_run: symbol " << name << " missing" << std::endl;
    return *out != nullptr;
}

bool load_api(const std::string& path, Api* a) {
    a->h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (!a->h) {
        std::cerr << "rpcf_run: cannot load " << path << ": " << dlerror() << std::endl;
        return false;
    }
    return sym(a->h, "rpcf_sizes", &a->sizes) && sym(a->h, "rpcf_create", &a->create) &&
           sym(a->h, "rpcf_destroy", &a->destroy) && sym(a->h, "rpcf_last_error", &a->last_error) &&
           sym(a->h, "rpcf_set_state", &a->set_state) && sym(a->h, "rpcf_get_state", &a->get_state) &&
           sym(a->h, "rpcf_get_nonlinear", &a->get_nonlinear) && sym(a->h, "rpcf_set_nonlinear", &a->set_nonlinear) &&
           sym(a->h, "rpcf_step", &a->step) && sym(a->h, "rpcf_bootstrap", &a->bootstrap) &&
           sym(a->h, "rpcf_seed_state", &a->seed_state) && sym(a->h, "rpcf_get_field", &a->get_field) &&
           sym(a->h, "rpcf_phys_to_spec", &a->phys_to_spec) && sym(a->h, "rpcf_spec_to_phys", &a->spec_to_phys);
}

struct Config {
    rpcf_params prm{};
    int start_type = 0, start_step = 0, end_step = 0, save_internal = 1, save_recovery_internal = 1;
    std::string output_prefix, recovery_prefix;
    bool seeded = false;  // line 8 present: seeded field; absent: the reference's analytic field
    uint64_t seed = 0;
    double amplitude = 1e-2, kexp = 0.0;
    int peak_bound = 0;
    int kmax = 4, degmax = 8;
};

// RPCF_Paras::read_para (source/operation.cpp:115-138) plus the two extension lines
bool read_para(const std::string& file, Config* c) {
    std::ifstream in(file);
    if (!in.is_open()) {
        std::cerr << "Error in opening file: " << file << std::endl;
        return false;
    }
    std::vector<std::string> lines;
    for (std::string l; std::getline(in, l);)
        if (l.find_first_not_of(" \t\r") != std::string::npos) lines.push_back(l);
    if (lines.size() < 6) {
        std::cerr << "rpcf_run: " << file << ": expected at least 6 parameter lines" << std::endl;
        return false;
    }
    std::istringstream s(lines[0] + " " + lines[1] + " " + lines[2] + " " + lines[3] + " " + lines[4] + " " + lines[5]);
    rpcf_params& p = c->prm;
    s >> p.mx >> p.my >> p.mz >> p.n_pi_x >> p.n_pi_y >> p.Re >> p.Ro >> p.dt >> c->start_type >> c->start_step >>
        c->end_step >> c->save_internal >> c->save_recovery_internal >> c->output_prefix >> c->recovery_prefix;
    if (s.fail()) {
        std::cerr << "rpcf_run: " << file << ": malformed parameter lines" << std::endl;
        return false;
    }
    p.u_profile[0] = 0.5;  // reference base flow U = (1+y)/2 (source/coefficient.cu:149-159)
    p.u_profile[1] = 0.5;
    p.u_profile[2] = 0.0;
    p.flow_control = 0;
    p.precision = 0;
    if (lines.size() > 6) {
        std::istringstream e(lines[6]);
        e >> p.u_profile[0] >> p.u_profile[1] >> p.u_profile[2] >> p.flow_control >> p.precision;
        if (e.fail()) {
            std::cerr << "rpcf_run: line 7 must be: c0 c1 c2 flow_control precision" << std::endl;
            return false;
        }
    }
    if (lines.size() > 7) {
        std::istringstream e(lines[7]);
        e >> c->seed >> c->amplitude >> c->kmax >> c->degmax >> c->kexp;
        if (e.fail()) {
            std::cerr << "rpcf_run: line 8 must be: seed amplitude kmax degmax kexp [peak_bound]" << std::endl;
            return false;
        }
        int pb = 0;
        if (e >> pb) c->peak_bound = pb;
        c->seeded = true;
    }
    return true;
}

struct Sizes {
    int nx, ny, nz, pz, hnx;
    size_t P;
};

// The reference's analytic initial field (init_flow_kernel and the _get_init_* profiles,
// source/RPCFKernels.cu:221-356, amplitude EPSILON_INIT = 0.005 at source/RPCFKernels.cu:132):
// wall-normal velocity w and wall-normal vorticity omega_z at the padded x/span points and
// the nz collocation planes z_p = cos(pi p/(nz-1)) (initFlow passes nz as the plane count,
// source/RPCFKernels.cu:372-376), zero on the two wall planes, then the unpadded forward
// transform (source/RPCFKernels.cu:401-406).  The (0,0) pencils -- the mean of w and
// omega_z, zero for this field -- become u0 and w0 as in initFlow (409-418).
struct InitProfile {
    double lx, ly;
    static constexpr double kEps = 0.005;
    double w(double x, double y, double z) const {
        const double pi = 3.14159265358979323846;
        const double cx1 = 2 * pi * x / lx, cx2 = 4 * pi * x / lx, cy1 = 2 * pi * y / ly, cy2 = 4 * pi * y / ly;
        return kEps * (-2.0 / 3.0) * (1.0 + cos(1.5 * pi * z)) *
               (sin(cx1) * sin(cy1) + sin(cx2) * sin(cy1) + sin(cx1) * sin(cy2));
    }
    double omega_z(double x, double y, double z) const {
        const double pi = 3.14159265358979323846;
        const double cx1 = 2 * pi * x / lx, cx2 = 4 * pi * x / lx, cy1 = 2 * pi * y / ly, cy2 = 4 * pi * y / ly;
        return kEps * 2.0 * pi * sin(1.5 * pi * z) *
               (lx / ly * (cos(cx1) * cos(cy1) + 0.5 * cos(cx2) * cos(cy1) + 2.0 * cos(cx1) * cos(cy2)) +
                ly / lx * (0.5 * cos(cx1) * cos(cy1) + cos(cx2) * cos(cy1) + 0.25 * cos(cx1) * cos(cy2)));
    }
};

bool reference_initial_state(const Api& a, rpcf_ctx* ctx, const rpcf_params& p, int nz, size_t P,
                             std::vector<double>* v, std::vector<double>* e) {
    const double pi = 3.14159265358979323846;
    // lx = 2 pi aphi, ly = 2 pi beta with aphi = n_pi_x/2, beta = n_pi_y/2 (source/data.cpp:5-6; include/data.h:321-322)
    const InitProfile prof{pi * p.n_pi_x, pi * p.n_pi_y};
    const size_t plane = (size_t)p.mx * p.my;
    std::vector<double> fw(plane * nz, 0.0), fo(plane * nz, 0.0);
    for (int k = 1; k < nz - 1; k++) {
        const double z = cos(pi * k / (nz - 1));
        for (int j = 0; j < p.my; j++) {
            const double y = (double)j / p.my * prof.ly;
            for (int i = 0; i < p.mx; i++) {
                const double x = (double)i / p.mx * prof.lx;
                fw[plane * k + (size_t)p.mx * j + i] = prof.w(x, y, z);
                fo[plane * k + (size_t)p.mx * j + i] = prof.omega_z(x, y, z);
            }
        }
    }
    v->assign(2 * P * nz, 0.0);
    e->assign(2 * P * nz, 0.0);
    return a.phys_to_spec(ctx, fw.data(), v->data(), nz, 0) == 0 && a.phys_to_spec(ctx, fo.data(), e->data(), nz, 0) == 0;
}

// output_velocity / write_velocity (source/output.cpp:18-103): the fields u, v, w (file
// <prefix>_velocity.<step>.dat) and omega_x, omega_y, omega_z (<prefix>_vortricity.<step>.dat,
// the reference's spelling) on the padded physical grid, raw doubles, field after field, each
// as pz planes of my rows of mx values.  Also prints the maximum magnitude of every field, as
// statistics() does (source/statistics.cpp:13-64).
bool write_velocity_files(const Api& a, rpcf_ctx* ctx, const Config& c, const Sizes& z, int step) {
    const rpcf_params& p = c.prm;
    const size_t n = 2 * z.P * z.nz, np = (size_t)z.pz * p.my * p.mx;
    std::vector<double> spec(n), phys(np);
    static const char* const kNames[6] = {"u", "v", "w", "ox", "oy", "oz"};
    std::ofstream out;
    for (int f = 0; f < 6; f++) {
        if (a.get_field(ctx, f, spec.data(), z.nz) || a.spec_to_phys(ctx, spec.data(), phys.data(), z.nz, 1))
            return false;
        double mx = 0.0;
        for (double x : phys) mx = std::max(mx, std::fabs(x));
        std::cout << "max " << kNames[f] << ":" << mx << std::endl;
        if (c.output_prefix.empty() || c.output_prefix == "-") continue;
        if (f % 3 == 0) {
            out.close();
            const std::string file = c.output_prefix + (f == 0 ? "_velocity." : "_vortricity.") + std::to_string(step) + ".dat";
            std::cout << "writing results to " << file << std::endl;
            out.open(file, std::ios::binary);
            if (!out) {
                std::cerr << "Cannot open output file:" << file << std::endl;
                return false;
            }
        }
        out.write((const char*)phys.data(), sizeof(double) * np);
        if (!out) return false;
    }
    return true;
}

const char kMagic[8] = {'R', 'P', 'C', 'F', 'R', 'C', 'V', '1'};

// write_recover_data (source/output.cpp:105-135) in a portable layout: magic, mx my mz nz,
// P, step, then v-hat, eta-hat, N_v, N_eta (pencil stride nz, complex double).  The
// (0,0) pencils carry u0, w0 and the (0,0) Lamb copies of the reference's header vectors.
bool write_recovery(const Api& a, rpcf_ctx* ctx, const Config& c, const Sizes& z, int step) {
    const size_t n = 2 * z.P * z.nz;
    std::vector<double> v(n), e(n), nv(n), ne(n);
    if (a.get_state(ctx, v.data(), e.data(), z.nz) || a.get_nonlinear(ctx, 0, nv.data(), ne.data(), z.nz)) return false;
    const std::string file = c.recovery_prefix + ".rcv";
    std::ofstream out(file, std::ios::binary);
    if (!out) {
        std::cerr << "cannot open recovery file " << file << std::endl;
        return false;
    }
    const int32_t hdr[4] = {c.prm.mx, c.prm.my, c.prm.mz, z.nz};
    const int64_t P = (int64_t)z.P, st = step;
    out.write(kMagic, 8);
    out.write((const char*)hdr, sizeof(hdr));
    out.write((const char*)&P, 8);
    out.write((const char*)&st, 8);
    for (auto* b : {&v, &e, &nv, &ne}) out.write((const char*)b->data(), sizeof(double) * n);
    return (bool)out;
}

// read_recover_data (source/output.cpp:141-178): the header must match (READ_AND_CHECK)
bool read_recovery(const Api& a, rpcf_ctx* ctx, const Config& c, const Sizes& z) {
    const std::string file = c.recovery_prefix + ".rcv";
    std::ifstream in(file, std::ios::binary);
    if (!in) {
        std::cerr << "cannot open recovery file " << file << std::endl;
        return false;
    }
    char magic[8];
    int32_t hdr[4];
    int64_t P, st;
    in.read(magic, 8);
    in.read((char*)hdr, sizeof(hdr));
    in.read((char*)&P, 8);
    in.read((char*)&st, 8);
    if (!in || memcmp(magic, kMagic, 8) || hdr[0] != c.prm.mx || hdr[1] != c.prm.my || hdr[2] != c.prm.mz ||
        hdr[3] != z.nz || P != (int64_t)z.P) {
        std::cerr << "wrong parameter" << std::endl;
        return false;
    }
    const size_t n = 2 * z.P * z.nz;
    std::vector<double> v(n), e(n), nv(n), ne(n);
    for (auto* b : {&v, &e, &nv, &ne}) in.read((char*)b->data(), sizeof(double) * n);
    if (!in) {
        std::cerr << "recovery file truncated" << std::endl;
        return false;
    }
    return a.set_state(ctx, v.data(), e.data(), z.nz) == 0 && a.set_nonlinear(ctx, 0, nv.data(), ne.data(), z.nz) == 0;
}

// spectral snapshot <output_prefix>_state.<step>.bin: v-hat then eta-hat
bool write_snapshot(const Api& a, rpcf_ctx* ctx, const Config& c, const Sizes& z, int step, double* u00) {
    const size_t n = 2 * z.P * z.nz;
    std::vector<double> v(n), e(n);
    if (a.get_state(ctx, v.data(), e.data(), z.nz)) return false;
    u00[0] = v[0];
    u00[1] = v[1];
    if (c.output_prefix.empty() || c.output_prefix == "-") return true;
    const std::string file = c.output_prefix + "_state." + std::to_string(step) + ".bin";
    std::ofstream out(file, std::ios::binary);
    if (!out) {
        std::cerr << "cannot open output file " << file << std::endl;
        return false;
    }
    out.write((const char*)v.data(), sizeof(double) * n);
    out.write((const char*)e.data(), sizeof(double) * n);
    return (bool)out;
}

std::string default_lib(const char* argv0) {
    std::string exe(argv0);
    const size_t slash = exe.rfind('/');
    return (slash == std::string::npos ? std::string(".") : exe.substr(0, slash)) + "/librpcf_hip.so";
}

}  // namespace

int main(int argc, char** argv) {
    std::string lib = default_lib(argv[0]);
    std::string para = "../input/parameter.txt";
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "--lib") && i + 1 < argc) lib = argv[++i];
        else para = argv[i];
    }
    std::cout << "Reading settings from " << para << std::endl;
    Config cfg;
    if (!read_para(para, &cfg)) return 2;
    Api api;
    if (!load_api(lib, &api)) return 2;
    Sizes z;
    if (api.sizes(&cfg.prm, &z.nx, &z.ny, &z.nz, &z.pz, &z.hnx)) {
        std::cerr << "rpcf_run: unsupported grid " << cfg.prm.mx << " " << cfg.prm.my << " " << cfg.prm.mz << std::endl;
        return 2;
    }
    z.P = (size_t)z.hnx * z.ny;
    rpcf_ctx* ctx = nullptr;
    std::cout << "init solver" << std::endl;
    if (api.create(&cfg.prm, &ctx)) {
        std::cerr << "rpcf_run: rpcf_create failed" << std::endl;
        return 1;
    }
    auto fail = [&](const char* what) {
        std::cerr << "rpcf_run: " << what << ": " << api.last_error(ctx) << std::endl;
        api.destroy(ctx);
        return 1;
    };
    double u00[2] = {0, 0};
    if (cfg.start_type == 0) {
        std::cout << "init flow" << std::endl;
        const size_t n = 2 * z.P * z.nz;
        std::vector<double> v(n), e(n);
        if (cfg.seeded) {
            if (api.seed_state(&cfg.prm, cfg.seed, cfg.amplitude, cfg.kmax, cfg.degmax, cfg.kexp, cfg.peak_bound,
                               v.data(), e.data(),
                               z.nz))
                return fail("seeded initial field");
        } else if (!reference_initial_state(api, ctx, cfg.prm, z.nz, z.P, &v, &e)) {
            return fail("analytic initial field");
        }
        if (api.set_state(ctx, v.data(), e.data(), z.nz)) return fail("set_state");
        std::cout << "output initial flow field" << std::endl;
        if (!write_snapshot(api, ctx, cfg, z, 0, u00)) return fail("snapshot");
        if (!write_velocity_files(api, ctx, cfg, z, 0)) return fail("velocity output");
        // extra previous step + re-initialisation + real first step (source/main_loop.cpp:39-52)
        std::cout << "first step" << std::endl;
        if (api.bootstrap(ctx)) return fail("bootstrap");
    } else if (cfg.start_type == 1) {
        if (!read_recovery(api, ctx, cfg, z)) {
            api.destroy(ctx);
            return 1;
        }
    } else {
        std::cerr << "wrong start type" << std::endl;
        api.destroy(ctx);
        return 2;
    }
    double total = 0.0;
    int count = 0;
    for (int i = cfg.start_step; i <= cfg.end_step; i++) {
        std::cout << std::endl << "STEP: " << i << std::endl;
        const auto t0 = std::chrono::steady_clock::now();
        if (api.step(ctx, 1)) return fail("step");
        const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        total += t;
        count++;
        std::cout << "Time cost of current step:" << t << std::endl;
        std::cout << "Mean time cost per step:" << total / count << std::endl;
        if (cfg.save_internal > 0 && i % cfg.save_internal == 0) {
            std::cout << "[OUTPUT] WRITING RESULTS" << std::endl;
            if (!write_snapshot(api, ctx, cfg, z, i, u00)) return fail("snapshot");
            if (!write_velocity_files(api, ctx, cfg, z, i)) return fail("velocity output");
            std::cout << "velocity(0,0,0):" << u00[0] << " " << u00[1] << std::endl;
        }
        if (cfg.save_recovery_internal > 0 && i % cfg.save_recovery_internal == 0 && i > 0) {
            std::cout << "[OUTPUT] WRITING RECOVERY FILES" << std::endl;
            if (!write_recovery(api, ctx, cfg, z, i)) return fail("recovery");
        }
    }
    api.destroy(ctx);
    return 0;
}
