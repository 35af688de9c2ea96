// 01_acoustic_scattering.cpp's run (01:497-547) on the GPU engine's legacy family-A mode:
// the quasi-2D cylinder scatterer (radius 0.5 on a 16 x 16 x 0.4 box at dx 0.05, z periodic),
// the initial Gaussian density pulse (01:166-205), probes A, B, C (01:157-163), frames
// output_mpi/cyl_<step>.vtk at step 0 and at t >= 3 and t >= 6, and the probe files
// output_mpi/probe_<name>.dat after (int)(11/dt) + 20 steps.
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nlos_engine.h"

namespace {

void check(int rc) {
    if (rc != NLOS_OK) throw std::runtime_error(std::string("nlos engine: ") + nlos_last_error());
}

// 01's Solver members (01:56-68) and its init_physics / run_simulation.
struct Solver {
    double X_MIN = -8.0, X_MAX = 8.0, Y_MIN = -8.0, Y_MAX = 8.0, Z_MIN = -0.2, Z_MAX = 0.2;
    double dx = 0.05, c = 1.0, tau = 0.505;
    double pulse_amp_lbm = 0.01;
    double output_scale_factor = 100.0;
    int nx = 0, ny = 0, nz = 0;
    std::vector<double> rho0;   // the initial density of every interior cell (1 on solids)
    nlos_engine *e = nullptr;

    ~Solver() { if (e) nlos_destroy(e); }

    void setup() {
        nlos_config cfg;
        check(nlos_config_default(&cfg));
        cfg.family = NLOS_FAMILY_A;
        cfg.precision = NLOS_F64;
        cfg.stream_pattern = NLOS_PATTERN_TWO_ARRAY;
        cfg.x_min = X_MIN; cfg.x_max = X_MAX; cfg.y_min = Y_MIN; cfg.y_max = Y_MAX;
        cfg.z_min = Z_MIN; cfg.z_max = Z_MAX;
        cfg.dx = dx;
        cfg.c_phys = c;
        cfg.tau = tau;
        cfg.sponge_width = 0.5;      // 01:172-173, 01's default sponge rule
        cfg.sponge_strength = 1.0;
        check(nlos_create(&cfg, &e));
        nlos_info info;
        check(nlos_get_info(e, &info));
        nx = info.nx; ny = info.ny; nz = info.nz;
    }

    // 01:166-205: solid where x^2 + y^2 < r^2; elsewhere rho = 1 + amp exp(-ln 2 d^2 / b^2).
    void init_physics() {
        const double cyl_radius = 0.5, cx = 4.0, cy = 0.0, b = 1.0;
        const size_t n = (size_t)nx * ny * nz;
        std::vector<uint8_t> solid(n, 0);
        rho0.assign(n, 1.0);
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const size_t c0 = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
                    const double x = X_MIN + (i + 0.5) * dx, y = Y_MIN + (j + 0.5) * dx;
                    if (x * x + y * y < cyl_radius * cyl_radius) {
                        solid[c0] = 1;
                    } else {
                        const double d2 = (x - cx) * (x - cx) + (y - cy) * (y - cy);
                        rho0[c0] = 1.0 + pulse_amp_lbm * std::exp(-std::log(2.0) * d2 / (b * b));
                    }
                }
        check(nlos_set_mask(e, solid.data()));
        check(nlos_set_density(e, rho0.data()));
        const double r135 = 135.0 * M_PI / 180.0;
        const double xyz[9] = {0.0, 1.0, 0.0, std::cos(r135), std::sin(r135), 0.0, -1.0, 0.0, 0.0};
        check(nlos_set_probes(e, xyz, 3, nullptr));
    }

    // The frame of step 0 is the initial density itself (01 writes rho before any step);
    // later frames come from the engine's captured field.
    void write_initial_vtk() {
        std::FILE *f = std::fopen("output_mpi/cyl_00000.vtk", "w");
        if (!f) throw std::runtime_error("cannot write output_mpi/cyl_00000.vtk");
        std::fprintf(f, "# vtk DataFile Version 3.0\nLBM Final Corrected\nASCII\nDATASET STRUCTURED_POINTS\n");
        std::fprintf(f, "DIMENSIONS %d %d %d\n", nx, ny, nz);
        std::fprintf(f, "ORIGIN %g %g %g\n", X_MIN, Y_MIN, Z_MIN);
        std::fprintf(f, "SPACING %g %g %g\n", dx, dx, dx);
        std::fprintf(f, "POINT_DATA %zu\n", rho0.size());
        std::fprintf(f, "SCALARS Pressure double 1\nLOOKUP_TABLE default\n");
        for (double v : rho0) std::fprintf(f, "%.6e\n", (v - 1.0) * output_scale_factor);
        std::fclose(f);
    }

    void run_simulation() {
        nlos_info info;
        check(nlos_get_info(e, &info));
        const double dt = info.dt;
        const int max_steps = (int)(11.0 / dt) + 20;
        const double vtk_times[2] = {3.0, 6.0};
        int next_vtk = 0;
        std::filesystem::create_directories("output_mpi");
        write_initial_vtk();
        int done = 0;   // steps taken; the state of step s is the one after s steps
        for (int s = 1; s <= max_steps; ++s) {
            const double time = s * dt;
            if (next_vtk < 2 && time >= vtk_times[next_vtk]) {
                ++next_vtk;
                check(nlos_step(e, s - done, NLOS_STEP_CAPTURE_FIELD));
                done = s;
                char fn[64];
                std::snprintf(fn, sizeof fn, "output_mpi/cyl_%05d.vtk", s);
                std::printf("Output at step %d t=%g\n", s, time);
                check(nlos_write_vtk(e, fn));
            }
        }
        // 01 records the states after s = 0 .. max_steps steps (each before its next step;
        // the step after the last record changes no output)
        check(nlos_step(e, max_steps - done, 0));
        const char *names[3] = {"A", "B", "C"};
        check(nlos_write_probe_dat(e, "output_mpi", names, output_scale_factor));
    }
};

}  // namespace

int main() {
    try {
        Solver s;
        s.setup();
        s.init_physics();
        s.run_simulation();
    } catch (const std::exception &ex) {
        std::fprintf(stderr, "%s\n", ex.what());
        return 1;
    }
    return 0;
}
