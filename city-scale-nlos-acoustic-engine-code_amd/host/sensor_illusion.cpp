// The configuration of 05_sensor_illusion.cpp's main (05:1124-1174) on the GPU engine.
// Reads csx_model_binary.stl from the working directory like the reference; writes
// probe_B1/B2/C1/C2.txt and, unless --no-vtk is given, output_mpi/flight_*.vtk.
// --dx <value> selects the spacing of 06 (1.0) or 07 (0.5); --precision f64|f32|f16s.
#include <cstring>
#include <string>

#include "lbm_solver.hpp"

int main(int argc, char **argv) {
    nlos_host::LBMSolver s;
    s.DX = 1.5;
    for (int a = 1; a < argc; ++a) {
        if (!std::strcmp(argv[a], "--no-vtk")) s.write_frames = false;
        else if (!std::strcmp(argv[a], "--dx") && a + 1 < argc) s.DX = std::stod(argv[++a]);
        else if (!std::strcmp(argv[a], "--precision") && a + 1 < argc) {
            const std::string p = argv[++a];
            s.precision = p == "f32" ? NLOS_F32 : (p == "f16s" ? NLOS_F16S : NLOS_F64);
        }
    }
    s.stl_file = "csx_model_binary.stl";
    s.stl_sx = -22248.0;
    s.stl_sy = -4123.0;
    s.stl_sz = -14.0;
    s.X_MIN = 668.0;
    s.X_MAX = 1064.0;
    s.Y_MIN = 17.0;
    s.Y_MAX = 314.0;
    s.Z_MIN = 0.0;
    s.Z_MAX = 106.0;
    const double B_loc_x = 23106.0 + s.stl_sx, B_loc_y = 4343.0 + s.stl_sy, B_loc_z = 100.0 + s.stl_sz;
    const double C_loc_x = 23026.0 + s.stl_sx, C_loc_y = 4313.0 + s.stl_sy, C_loc_z = 100.0 + s.stl_sz;
    s.probes.push_back({"B1", B_loc_x, B_loc_y, B_loc_z, {}});
    s.probes.push_back({"B2", B_loc_x + s.DX, B_loc_y, B_loc_z, {}});
    s.probes.push_back({"C1", C_loc_x, C_loc_y, C_loc_z, {}});
    s.probes.push_back({"C2", C_loc_x + s.DX, C_loc_y, C_loc_z, {}});
    s.drone.sx = 912.0; s.drone.sy = 66.6; s.drone.sz = 86.0;
    s.drone.ex = 912.0; s.drone.ey = 66.6; s.drone.ez = 86.0;
    s.drone.speed = 0.0;
    s.drone.amp = 10.0;
    s.drone.freq = 0.0;
    s.output_world_coords = true;
    try {
        s.init(argc, argv);
        s.run(1, 80);
    } catch (const std::exception &ex) {
        std::fprintf(stderr, "%s\n", ex.what());
        return 1;
    }
    return 0;
}
