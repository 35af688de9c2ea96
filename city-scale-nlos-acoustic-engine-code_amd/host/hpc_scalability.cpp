// 04_hpc_scalability.cpp's benchmark (04:360-404) on the GPU engine's legacy family-A mode:
// the +-13 m box at dx 0.125 (208^3 cells), tau 0.51, the solid cylinder r < 4 along z, the
// radial sponge beyond r = 10 over 3 m (strength 0.5), 04's line source (04:212-235); 100
// warm-up steps, then 2000 timed steps with the source time counted from 0 again, and
// 04's report: BENCHMARK_TIME and MLUPS over all nodes (04:399-401).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nlos_engine.h"

namespace {

void check(int rc) {
    if (rc != NLOS_OK) throw std::runtime_error(std::string("nlos engine: ") + nlos_last_error());
}

}  // namespace

int main() {
    const double X_MIN = -13.0, X_MAX = 13.0, dx = 0.125, c = 1.0, tau = 0.51;
    const double cyl_R = 4.0, sponge_start_R = 10.0, sponge_thickness = 3.0, damping_strength = 0.5;
    const int warmup = 100, fixed_steps = 2000;
    nlos_engine *e = nullptr;
    try {
        nlos_config cfg;
        check(nlos_config_default(&cfg));
        cfg.family = NLOS_FAMILY_A;
        cfg.precision = NLOS_F64;
        cfg.stream_pattern = NLOS_PATTERN_TWO_ARRAY;
        cfg.x_min = cfg.y_min = cfg.z_min = X_MIN;
        cfg.x_max = cfg.y_max = cfg.z_max = X_MAX;
        cfg.dx = dx;
        cfg.c_phys = c;
        cfg.tau = tau;
        check(nlos_create(&cfg, &e));
        nlos_info info;
        check(nlos_get_info(e, &info));
        const int nx = info.nx, ny = info.ny, nz = info.nz;
        // init_physics (04:176-208): 04's cell centres x = X_MIN + i dx + dx/2
        const size_t n = (size_t)nx * ny * nz;
        std::vector<uint8_t> solid(n, 0);
        std::vector<double> sigma(n, 0.0);
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const size_t c0 = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
                    const double x = X_MIN + (double)i * dx + 0.5 * dx, y = X_MIN + (double)j * dx + 0.5 * dx;
                    const double r = std::sqrt(x * x + y * y);
                    if (r < cyl_R) {
                        solid[c0] = 1;
                    } else if (r > sponge_start_R) {
                        double ratio = (r - sponge_start_R) / sponge_thickness;
                        if (ratio > 1.0) ratio = 1.0;
                        sigma[c0] = damping_strength * ratio * ratio;
                    }
                }
        check(nlos_set_mask(e, solid.data()));
        check(nlos_set_sponge(e, sigma.data()));
        nlos_line_source src{0.0, -6.0, 0.01, 2.0, M_PI / 15.0, 36.0, 0.577350269 * dx / c, 0};
        check(nlos_set_line_source(e, &src));

        std::printf("=== START BENCHMARK ===\nGrid: %d x %d x %d (dx=%g)\nSteps=%d, GPUs=1\n", nx, ny, nz, dx,
                    fixed_steps);
        check(nlos_step(e, warmup, 0));
        src.step0 = warmup;   // the timed loop counts t from 0 again (04:381-388)
        check(nlos_set_line_source(e, &src));
        check(nlos_synchronize(e));
        const auto t0 = std::chrono::high_resolution_clock::now();
        check(nlos_step(e, fixed_steps, 0));
        check(nlos_synchronize(e));
        const auto t1 = std::chrono::high_resolution_clock::now();
        const double total_time = std::chrono::duration<double>(t1 - t0).count();
        const double total_nodes = (double)nx * (double)ny * (double)nz;
        std::printf("BENCHMARK_TIME: %g s\nMLUPS: %g\n=== END BENCHMARK ===\n", total_time,
                    total_nodes * fixed_steps / total_time / 1e6);
        nlos_destroy(e);
    } catch (const std::exception &ex) {
        std::fprintf(stderr, "%s\n", ex.what());
        if (e) nlos_destroy(e);
        return 1;
    }
    return 0;
}
