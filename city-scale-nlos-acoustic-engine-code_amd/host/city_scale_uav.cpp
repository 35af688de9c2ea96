// The configuration of 08_city_scale_uav.cpp's main (08:1062-1092) on the GPU engine: the
// member defaults of the box (0..1650 x 0..1150 x 0..250 m), DX = 1.5, the STL shift, a 20 Hz
// sine source of amplitude 0.5 flying along +y at 30 m/s (08:540-628), run(3.0, 3).
// Reads csx_model_binary.stl from the working directory like the reference and writes
// output_mpi/flight_*.vtk.  Options: --binary (legacy VTK BINARY frames), --max-frames N
// (stop after N frames), --no-vtk, --precision f64|f32|f16s|b16.
#include <cstring>
#include <string>

#include "lbm_solver.hpp"

int main(int argc, char **argv) {
    nlos_host::LBMSolver s;
    s.DX = 1.5;
    for (int a = 1; a < argc; ++a) {
        if (!std::strcmp(argv[a], "--no-vtk")) s.write_frames = false;
        else if (!std::strcmp(argv[a], "--binary")) s.binary_frames = true;
        else if (!std::strcmp(argv[a], "--max-frames") && a + 1 < argc) s.max_frames = std::stoi(argv[++a]);
        else if (!std::strcmp(argv[a], "--precision") && a + 1 < argc) {
            const std::string p = argv[++a];
            s.precision = p == "f32" ? NLOS_F32 : p == "f16s" ? NLOS_F16S : p == "b16" ? NLOS_B16 : NLOS_F64;
        }
    }
    s.stl_file = "csx_model_binary.stl";
    s.stl_sx = -22248.0;
    s.stl_sy = -4123.0;
    s.stl_sz = -14.0;
    s.drone.sx = 912.0; s.drone.sy = 66.6; s.drone.sz = 86.0;
    s.drone.ex = 912.0; s.drone.ey = 1127.3; s.drone.ez = 86.0;
    s.drone.speed = 30.0;
    s.drone.amp = 0.5;
    s.drone.freq = 20.0;
    s.waveform = NLOS_WAVE_SINE;   // 08:543
    s.output_world_coords = true;
    try {
        s.init(argc, argv);
        s.run(3.0, 3);
    } catch (const std::exception &ex) {
        std::fprintf(stderr, "%s\n", ex.what());
        return 1;
    }
    return 0;
}
