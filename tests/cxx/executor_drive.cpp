// A C++ caller of the library, built with g++ against include/glisser_amd.h
// and include/glisser_amd_executor.hpp and linked to libglisser_amd.so, as a
// GLISSER host would be (the reference host is C++: src/executor.cu,
// src/executor_facade.h).  Test driver for tests/test_gpu_cxx_caller.py.
//
//   executor_drive abi  <input> <output>   the reference loop's device calls,
//                                          one by one (Executor::init, loop,
//                                          resync2 replaced call by call)
//   executor_drive exec <input> <output>   the library's own loop, glx::Executor
//
// Input (little-endian, written by the test): u64 n, u32 n_bodies, u32
// tbsize, u32 n_blocks, u32 form, f64 dt, f64 outer, f64 cull, f64 rh_factor,
// then m[nb], r[nb][3], v[nb][3], pr[n][3], pv[n][3], id[n] (u32), h0 at t0
// [3], rh at t0 [nb], and per block rh[nb], r_log[tbsize][nb-1][3],
// h0_log[tbsize][3] (the host planet producer's blocks).
// Output: u64 n_alive, r[n][3], v[n][3], flags[n] (u16), id[n] (u32),
// deathtime[n] (f64; NaN in abi mode), slot order.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "glisser_amd.h"
#include "glisser_amd_executor.hpp"

namespace {

// The wrapper a reference-side binding adds: status codes back to the
// reference's exceptions (src/util.h:10).
void glx_check(int status, const char* what) {
    if (status != GLX_OK) throw std::runtime_error(std::string(what) + ": " + glx_last_error());
}

struct Reader {
    std::ifstream f;
    explicit Reader(const char* path) : f(path, std::ios::binary) {
        if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    }
    template <typename T>
    T one() {
        T x;
        f.read(reinterpret_cast<char*>(&x), sizeof x);
        if (!f) throw std::runtime_error("input cut short");
        return x;
    }
    template <typename T>
    std::vector<T> many(size_t n) {
        std::vector<T> v(n);
        f.read(reinterpret_cast<char*>(v.data()), n * sizeof(T));
        if (!f) throw std::runtime_error("input cut short");
        return v;
    }
};

struct Input {
    uint64_t n;
    uint32_t nb, tb, n_blocks, form;
    double dt, outer, cull, rh_factor;
    std::vector<double> m, r, v, pr, pv, h0_t0, rh_t0;
    std::vector<uint32_t> id;
    std::vector<std::vector<double>> rh, r_log, h0_log;
};

Input read_input(const char* path) {
    Reader in(path);
    Input x;
    x.n = in.one<uint64_t>();
    x.nb = in.one<uint32_t>();
    x.tb = in.one<uint32_t>();
    x.n_blocks = in.one<uint32_t>();
    x.form = in.one<uint32_t>();
    x.dt = in.one<double>();
    x.outer = in.one<double>();
    x.cull = in.one<double>();
    x.rh_factor = in.one<double>();
    x.m = in.many<double>(x.nb);
    x.r = in.many<double>(3 * (size_t)x.nb);
    x.v = in.many<double>(3 * (size_t)x.nb);
    x.pr = in.many<double>(3 * x.n);
    x.pv = in.many<double>(3 * x.n);
    x.id = in.many<uint32_t>(x.n);
    x.h0_t0 = in.many<double>(3);
    x.rh_t0 = in.many<double>(x.nb);
    for (uint32_t k = 0; k < x.n_blocks; k++) {
        x.rh.push_back(in.many<double>(x.nb));
        x.r_log.push_back(in.many<double>((size_t)x.tb * (x.nb - 1) * 3));
        x.h0_log.push_back(in.many<double>((size_t)x.tb * 3));
    }
    return x;
}

void write_output(const char* path, uint64_t n_alive, const std::vector<double>& r, const std::vector<double>& v,
                  const std::vector<uint16_t>& flags, const std::vector<uint32_t>& id,
                  const std::vector<double>& deathtime) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(&n_alive), sizeof n_alive);
    f.write(reinterpret_cast<const char*>(r.data()), r.size() * sizeof(double));
    f.write(reinterpret_cast<const char*>(v.data()), v.size() * sizeof(double));
    f.write(reinterpret_cast<const char*>(flags.data()), flags.size() * sizeof(uint16_t));
    f.write(reinterpret_cast<const char*>(id.data()), id.size() * sizeof(uint32_t));
    f.write(reinterpret_cast<const char*>(deathtime.data()), deathtime.size() * sizeof(double));
    if (!f) throw std::runtime_error(std::string("cannot write ") + path);
}

// The reference loop in kill mode with the device half replaced call by
// call: Executor::init's allocations and upload,
// the initial accelerations of the WHIntegrator constructor, then per block
// the planet-log upload, the time-block kernel, KillEncounterKernel and the
// partition, and at the end download_data.
int run_abi(const Input& x, const char* out) {
    glx_params p{x.n, x.nb, x.tb, 0, 0, 2};
    glx_ctx* ctx = nullptr;
    glx_check(glx_create(&p, &ctx), "glx_create");
    std::vector<double> a(3 * x.n, 0.0);
    std::vector<uint16_t> flags(x.n, 0);
    std::vector<uint32_t> dti(x.n, 0);
    std::vector<uint32_t> pid(x.nb);
    for (uint32_t j = 0; j < x.nb; j++) pid[j] = j;
    glx_check(glx_upload_particles(ctx, 0, x.n, x.pr.data(), x.pv.data(), a.data(), flags.data(), x.id.data(),
                                   dti.data()),
              "glx_upload_particles");
    glx_check(glx_set_counts(ctx, x.n, x.n), "glx_set_counts");
    // helio_acc_particles<true> of the WHIntegrator constructor (src/wh.cpp:531-543)
    glx_check(glx_init_accel(ctx, x.nb, x.m.data(), pid.data(), x.rh_t0.data(), x.r.data() + 3, x.h0_t0.data(),
                             x.outer),
              "glx_init_accel");
    glx_check(glx_kill_encounters(ctx), "glx_kill_encounters");
    uint64_t n_alive = 0, n_enc = 0;
    glx_check(glx_partition(ctx, GLX_PART_UNFLAGGED, &n_alive, &n_enc), "glx_partition");
    for (uint32_t k = 0; k < x.n_blocks; k++) {
        const int buf = (int)(k % 2);  // dd.planet_data_id % 2
        glx_check(glx_upload_planet_block(ctx, buf, x.nb, x.m.data(), pid.data(), x.rh[k].data(), x.r_log[k].data(),
                                          x.h0_log[k].data(), x.tb),
                  "glx_upload_planet_block");
        glx_check(glx_run_block(ctx, buf, x.dt, x.outer, (int)x.form), "glx_run_block");
        glx_check(glx_kill_encounters(ctx), "glx_kill_encounters");
        glx_check(glx_partition(ctx, GLX_PART_UNFLAGGED, &n_alive, &n_enc), "glx_partition");
    }
    std::vector<double> r(3 * x.n), v(3 * x.n);
    std::vector<uint32_t> id(x.n);
    glx_check(glx_download_particles(ctx, 0, x.n, r.data(), v.data(), nullptr, flags.data(), id.data(), nullptr),
              "glx_download_particles");
    glx_check(glx_destroy(ctx), "glx_destroy");
    write_output(out, n_alive, r, v, flags, id, std::vector<double>(x.n, std::nan("")));
    return 0;
}

// The library's own loop: planets integrated on
// the host, kill mode.
int run_exec(const Input& x, const char* out) {
    glx::ExecutorConfig cfg;
    cfg.dt = x.dt;
    cfg.tbsize = x.tb;
    cfg.outer_radius = x.outer;
    cfg.cull_radius = x.cull;
    cfg.rh_factor = x.rh_factor;
    cfg.form = (int)x.form;
    glx::Executor ex(cfg, x.nb, x.m.data(), x.r.data(), x.v.data(), nullptr, x.n, x.pr.data(), x.pv.data(),
                     x.id.data());
    ex.run(x.n_blocks);
    const glx::ParticleState s = ex.download();
    write_output(out, s.n_alive, s.r, s.v, s.flags, s.id, s.deathtime);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s abi|exec <input> <output>\n", argv[0]);
        return 2;
    }
    try {
        const Input x = read_input(argv[2]);
        if (std::strcmp(argv[1], "abi") == 0) return run_abi(x, argv[3]);
        if (std::strcmp(argv[1], "exec") == 0) return run_exec(x, argv[3]);
        std::fprintf(stderr, "unknown mode %s\n", argv[1]);
        return 2;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
