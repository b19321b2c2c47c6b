// rpcf_run -- the simulation driver: the reference's run_simulation
// (source/main_loop.cpp:15-115) restated over the rpcf C ABI (include/rpcf.h).
//
//   rpcf_run [--lib PATH] [PARAMETER_FILE]
//
// PARAMETER_FILE (default ../input/parameter.txt, as the reference) holds the reference's
// six whitespace-separated lines (RPCF_Paras::read_para, source/operation.cpp:152-175):
//   mx my mz / n_pi_x n_pi_y / Re Ro dt /
//   start_type start_step end_step save_internal save_recovery_internal /
//   output_prefix / recovery_prefix
// and optionally two extension lines (absent: reference behaviour):
//   c0 c1 c2 flow_control precision    base flow U = c0 + c1 y + c2 y^2, flow-rate control, 0 FP64 / 1 mixed
//   seed amplitude kmax degmax kexp    seeded initial field (rpcf_seed_state)
//
// --lib selects the library implementing the ABI (default: librpcf_hip.so beside this
// executable); the same driver runs the CPU oracle for testing.
//
// Differences from the reference, by design: errors are reported and end the program
// with a non-zero status instead of exit()/system("pause"); the initial field is the
// seeded one (the reference's analytic init_flow_kernel is not reproduced); snapshots
// are spectral state files; the recovery file is portable (sizes in the header, pencil
// stride nz) and lives at <recovery_prefix>.rcv.
#include <dlfcn.h>

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
    int (*seed_state)(const rpcf_params*, uint64_t, double, int, int, double, double*, double*, long);
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
           sym(a->h, "rpcf_seed_state", &a->seed_state);
}

struct Config {
    rpcf_params prm{};
    int start_type = 0, start_step = 0, end_step = 0, save_internal = 1, save_recovery_internal = 1;
    std::string output_prefix, recovery_prefix;
    uint64_t seed = 0;
    double amplitude = 1e-2, kexp = 0.0;
    int kmax = 4, degmax = 8;
};

// RPCF_Paras::read_para (source/operation.cpp:152-175) plus the two extension lines
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
            std::cerr << "rpcf_run: line 8 must be: seed amplitude kmax degmax kexp" << std::endl;
            return false;
        }
    }
    return true;
}

struct Sizes {
    int nx, ny, nz, pz, hnx;
    size_t P;
};

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
        if (api.seed_state(&cfg.prm, cfg.seed, cfg.amplitude, cfg.kmax, cfg.degmax, cfg.kexp, v.data(), e.data(), z.nz))
            return fail("seeded initial field");
        if (api.set_state(ctx, v.data(), e.data(), z.nz)) return fail("set_state");
        std::cout << "output initial flow field" << std::endl;
        if (!write_snapshot(api, ctx, cfg, z, 0, u00)) return fail("snapshot");
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
