// lbm_solver.hpp - C++ host class with the surface of the reference's LBMSolver
// (05_sensor_illusion.cpp:241-1122), driving the GPU engine through its C ABI.
//
// The public fields keep the reference's names and meaning (05:243-268, 669), and the
// public methods init / apply_source / exchange_halos / step_collision_bgk / write_vtk /
// run behave like theirs: same grid rule, same time loop s = 0..round(T/dt), probes
// recorded after every step and written as probe_<name>.txt with "%.8e" lines, VTK frames
// output_mpi/flight_<frame>.vtk every max(1, total/Frames) steps.  A caller of the
// reference switches by including this header and linking libnlos_engine.so.
//
// Fields beyond the reference's constants (all default to the reference's values):
// c_phys, sponge width/strength/faces, sky_z, NSCBC faces, ground, wall reflection,
// precision, waveform (pulse of 05 or sine of 08_city_scale_uav.cpp:543), injection,
// write_frames, binary_frames, max_frames.
#pragma once

#include <algorithm>
#include <cmath>
#include <filesystem>
#include <cstdio>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nlos_engine.h"

namespace nlos_host {

// DroneSource (05:211-239): straight path start -> end at `speed`.
struct DroneSource {
    double sx = 0, sy = 0, sz = 0;
    double ex = 0, ey = 0, ez = 0;
    double speed = 0;
    double amp = 0;
    double freq = 0;
};

class LBMSolver {
public:
    std::string stl_file;
    double stl_sx = 0.0, stl_sy = 0.0, stl_sz = 0.0;
    DroneSource drone;
    bool use_geometry = true;
    double X_MIN = 0.0, X_MAX = 1650.0;
    double Y_MIN = 0.0, Y_MAX = 1150.0;
    double Z_MIN = 0.0, Z_MAX = 250.0;
    double DX = 2.0;
    bool output_world_coords = true;

    struct Probe {
        std::string name;
        double x, y, z;
        std::vector<double> p_history;
    };
    std::vector<Probe> probes;

    double tau = 0.55;
    bool use_nscbc = true;

    // Engine parameters that the reference hard-codes.
    double c_phys = 340.0;            // 05:198
    double sponge_cells = 20.0;       // sponge width in cells, 20*dx (05:455)
    double sponge_strength = 0.35;    // 05:456
    double sky_z = 120.0;             // 05:458
    int sponge_faces = NLOS_FACES_REFERENCE;
    int nscbc_faces = NLOS_FACES_REFERENCE;
    bool ground = true;
    double wall_reflect = 1.0;
    int precision = NLOS_F64;
    int voxelizer = NLOS_VOXEL_HEIGHT;   // the reference's height field; see nlos_voxelizer
    int waveform = NLOS_WAVE_GAUSS_PULSE;
    int injection = NLOS_INJECT_GAUSS5;
    bool write_frames = true;
    bool async_frames = true;         // frames written off the step stream (same files)
    bool binary_frames = false;       // legacy VTK BINARY instead of the reference's ASCII
    int max_frames = -1;              // >= 0: run stops once this many frames are written

    LBMSolver() = default;
    LBMSolver(const LBMSolver &) = delete;
    LBMSolver &operator=(const LBMSolver &) = delete;
    ~LBMSolver() { if (eng_) nlos_destroy(eng_); }

    // LBMSolver::init (05:299-450).  argc/argv are accepted and ignored, as there.
    void init(int argc, char **argv) {
        (void)argc;
        (void)argv;
        nlos_config c;
        check(nlos_config_default(&c));
        c.x_min = X_MIN; c.x_max = X_MAX; c.y_min = Y_MIN; c.y_max = Y_MAX; c.z_min = Z_MIN; c.z_max = Z_MAX;
        c.dx = DX;
        c.c_phys = c_phys;
        c.tau = tau;
        c.sponge_width = sponge_cells * DX;
        c.sponge_strength = sponge_strength;
        c.sky_z = sky_z;
        c.wall_reflect = wall_reflect;
        c.world_offset[0] = -stl_sx; c.world_offset[1] = -stl_sy; c.world_offset[2] = -stl_sz;  // 05:293-296
        c.sponge_faces = sponge_faces;
        c.nscbc_faces = use_nscbc ? nscbc_faces : 0;
        c.ground = ground ? 1 : 0;
        c.use_geometry = use_geometry ? 1 : 0;
        c.output_world_coords = output_world_coords ? 1 : 0;
        c.precision = precision;
        c.voxelizer = voxelizer;
        check(nlos_create(&c, &eng_));
        nlos_info info;
        check(nlos_get_info(eng_, &info));
        dt_ = info.dt;
        if (use_geometry) {
            const int rc = nlos_load_stl(eng_, stl_file.c_str(), 1.0, stl_sx, stl_sy, stl_sz);
            if (rc == NLOS_OK) std::cout << "STL loaded: " << stl_file << "\n";
            else if (rc == NLOS_ERR_IO) std::cout << "Warning: STL not loaded, using empty geometry.\n";
            else check(rc);
        }
        nlos_source s;
        check(nlos_source_default(&s));
        s.sx = drone.sx; s.sy = drone.sy; s.sz = drone.sz;
        s.ex = drone.ex; s.ey = drone.ey; s.ez = drone.ez;
        s.speed = drone.speed; s.amp = drone.amp; s.freq = drone.freq;
        s.waveform = waveform;
        s.injection = injection;
        check(nlos_set_source(eng_, &s));
        std::vector<double> xyz;
        for (auto &p : probes) { xyz.push_back(p.x); xyz.push_back(p.y); xyz.push_back(p.z); }
        std::vector<uint8_t> in_solid(probes.size(), 0);
        check(nlos_set_probes(eng_, xyz.data(), (int32_t)probes.size(), in_solid.data()));
        for (size_t p = 0; p < probes.size(); ++p)
            if (in_solid[p]) std::cout << "WARNING: Probe " << probes[p].name << " is trapped inside SOLID building!\n";
        std::cout << "\n--- D3Q19-BGK acoustic engine (GPU) ---\n";
        std::cout << "Grid: " << info.nx << " x " << info.ny << " x " << info.nz << "  (dx=" << DX << " m)\n";
        std::cout << "dt=" << dt_ << " s\n";
    }

    void apply_source(double t) { check(nlos_apply_source(eng_, t)); }     // 05:576-667
    void exchange_halos() { check(nlos_exchange_halos(eng_)); }            // 05:671-904
    void step_collision_bgk() { check(nlos_step_collision(eng_)); }        // 05:906-972

    // LBMSolver::write_vtk (05:975-1062): output_mpi/flight_<frame>.vtk of the field of
    // the last captured step.
    void write_vtk(int frame_id) {
        std::cout << "Write VTK frame " << frame_id << " ...\n";
        std::filesystem::create_directories("output_mpi");
        char fn[64];
        std::snprintf(fn, sizeof fn, "output_mpi/flight_%05d.vtk", frame_id);
        const int fmt = binary_frames ? NLOS_VTK_BINARY : NLOS_VTK_ASCII;
        if (async_frames) check(nlos_dump_field_async(eng_, fn, fmt));
        else check(binary_frames ? nlos_write_vtk_binary(eng_, fn) : nlos_write_vtk(eng_, fn));
    }

    // LBMSolver::run (05:1064-1121).
    void run(double T_total, int Frames) {
        const int total_steps = (int)std::round(T_total / dt_);
        const int out_interval = std::max(1, total_steps / Frames);
        std::cout << "Run: T_total=" << T_total << " s, Frames=" << Frames << ", total_steps=" << total_steps
                  << ", out_interval=" << out_interval << "\n";
        int frame_id = 0;
        int s = 0;
        while (s <= total_steps) {
            // advance to the next frame step (or to the end) in one fused call
            int next = ((s + out_interval - 1) / out_interval) * out_interval;
            const bool frame = write_frames && next <= total_steps;
            if (!frame) next = total_steps;
            check(nlos_step(eng_, next - s + 1, frame ? NLOS_STEP_CAPTURE_FIELD : 0));
            s = next + 1;
            if (frame) write_vtk(frame_id++);
            if (max_frames >= 0 && frame_id >= max_frames) break;
        }
        check(nlos_dump_wait(eng_));
        collect_probes();
        for (auto &p : probes) {
            FILE *f = std::fopen(("probe_" + p.name + ".txt").c_str(), "w");
            if (!f) throw std::runtime_error("cannot write probe file");
            for (double v : p.p_history) std::fprintf(f, "%.8e\n", v);
            std::fclose(f);
            std::cout << "successfully wrote probe_" << p.name << ".txt (" << p.p_history.size() << " steps)\n";
        }
        std::cout << "\nDone.\n";
    }

    // Copies the engine's probe histories into Probe::p_history (05:1085-1093).
    void collect_probes() {
        nlos_info info;
        check(nlos_get_info(eng_, &info));
        const long long ns = info.step;
        std::vector<double> buf(probes.size() * (size_t)std::max<long long>(ns, 1));
        int64_t got = 0;
        check(nlos_read_probes(eng_, buf.data(), ns, &got));
        for (size_t p = 0; p < probes.size(); ++p)
            probes[p].p_history.assign(buf.begin() + p * ns, buf.begin() + p * ns + got);
    }

    nlos_engine *engine() { return eng_; }

private:
    nlos_engine *eng_ = nullptr;
    double dt_ = 0.0;

    static void check(int rc) {
        if (rc != NLOS_OK) throw std::runtime_error(std::string("nlos engine: ") + nlos_last_error());
    }
};

}  // namespace nlos_host
