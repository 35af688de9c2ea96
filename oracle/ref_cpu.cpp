/*
 * ORACLE - test infrastructure only.  CPU restatement of the reference's family-B solver
 * (05_sensor_illusion.cpp, with the sine waveform of 08_city_scale_uav.cpp:543), in FP64,
 * array-of-structures, two arrays, OpenMP over cells, following the reference operation
 * for operation so that results agree bit for bit at one MPI rank.  It is the checker for
 * the HIP engine and the CPU baseline of bench.py; the engine never links it.
 *
 * Generalised only where the engine's configuration is wider than the reference's
 * hard-coded constants: sponge width/strength/face set, NSCBC face set, ground on/off,
 * sky height, wall reflection, trilinear injection, sampled and Ricker waveforms, initial
 * density.  With the reference's values every generalised branch reduces to the
 * reference's code path.
 *
 * Pins (tests/test_oracle.py): the compiled reference's probe files and VTK samples for
 * the flat and city scenes (tests/golden/ref05_*), closed forms of the lattice (weights,
 * inverse directions, equilibrium moments), the rest-state fixed point, mass
 * conservation in a closed box, the source normalisation, analytic box voxelisation.
 * Parity unpinned: NLOS_WAVE_SAMPLED/RICKER and NLOS_INJECT_TRILINEAR have no
 * reference counterpart; they are pinned only by closed-form checks of their weights.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "nlos_engine.h"

namespace {

constexpr int Q = 19;
constexpr double cs2 = 1.0 / 3.0;                    // 05:17
constexpr double cs = 0.57735026918962576451;         // 05:18
constexpr double PI = 3.14159265358979323846;         // 05:19

// Velocity ordering B (05:21-23) and weights (05:25-31).
constexpr int ex[Q] = {0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1, 0, 0, 0, 0};
constexpr int ey[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1};
constexpr int ez[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1};
constexpr double w[Q] = {
    1.0 / 3.0,
    1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};

enum : uint8_t { FLUID = 0, SOLID = 1 };   // 05:36

struct Point { double x, y, z; };
struct Triangle { Point v1, v2, v3, minPt, maxPt; };

#pragma pack(push, 1)
struct StlTriangle { float n[3]; float v1[3]; float v2[3]; float v3[3]; uint16_t attr; };
#pragma pack(pop)

// GeometryEngine (05:60-193).
struct Geometry {
    std::vector<Triangle> triangles;
    Point geoMin{1e30, 1e30, 1e30}, geoMax{-1e30, -1e30, -1e30};
    int gridRows = 200, gridCols = 200;
    double cellW = 1.0, cellH = 1.0;
    std::vector<std::vector<int>> grid;

    // 05:70-111, reading from memory instead of an ifstream.
    bool load(const unsigned char *buf, size_t len, double scale, double sx, double sy, double sz) {
        if (len < 84) return false;
        uint32_t numTri = 0;
        std::memcpy(&numTri, buf + 80, 4);
        if (len < 84 + (size_t)numTri * 50) return false;
        triangles.resize(numTri);
        geoMin = {1e30, 1e30, 1e30};
        geoMax = {-1e30, -1e30, -1e30};
        for (uint32_t i = 0; i < numTri; ++i) {
            StlTriangle st{};
            std::memcpy(&st, buf + 84 + (size_t)i * 50, 50);
            Triangle &t = triangles[i];
            t.v1 = {st.v1[0] * scale + sx, st.v1[1] * scale + sy, st.v1[2] * scale + sz};
            t.v2 = {st.v2[0] * scale + sx, st.v2[1] * scale + sy, st.v2[2] * scale + sz};
            t.v3 = {st.v3[0] * scale + sx, st.v3[1] * scale + sy, st.v3[2] * scale + sz};
            t.minPt.x = std::min({t.v1.x, t.v2.x, t.v3.x});
            t.minPt.y = std::min({t.v1.y, t.v2.y, t.v3.y});
            t.minPt.z = std::min({t.v1.z, t.v2.z, t.v3.z});
            t.maxPt.x = std::max({t.v1.x, t.v2.x, t.v3.x});
            t.maxPt.y = std::max({t.v1.y, t.v2.y, t.v3.y});
            t.maxPt.z = std::max({t.v1.z, t.v2.z, t.v3.z});
            geoMin.x = std::min(geoMin.x, t.minPt.x);
            geoMin.y = std::min(geoMin.y, t.minPt.y);
            geoMin.z = std::min(geoMin.z, t.minPt.z);
            geoMax.x = std::max(geoMax.x, t.maxPt.x);
            geoMax.y = std::max(geoMax.y, t.maxPt.y);
            geoMax.z = std::max(geoMax.z, t.maxPt.z);
        }
        buildGrid();
        return true;
    }

    // 05:113-142
    void buildGrid() {
        if (triangles.empty()) return;
        cellW = (geoMax.x - geoMin.x) / gridRows;
        cellH = (geoMax.y - geoMin.y) / gridCols;
        if (cellW < 1e-9) cellW = 1.0;
        if (cellH < 1e-9) cellH = 1.0;
        grid.assign(gridRows * gridCols, {});
        for (int i = 0; i < (int)triangles.size(); ++i) {
            const auto &t = triangles[i];
            int r1 = (int)((t.minPt.x - geoMin.x) / cellW);
            int r2 = (int)((t.maxPt.x - geoMin.x) / cellW);
            int c1 = (int)((t.minPt.y - geoMin.y) / cellH);
            int c2 = (int)((t.maxPt.y - geoMin.y) / cellH);
            r1 = std::max(0, std::min(gridRows - 1, r1));
            r2 = std::max(0, std::min(gridRows - 1, r2));
            c1 = std::max(0, std::min(gridCols - 1, c1));
            c2 = std::max(0, std::min(gridCols - 1, c2));
            for (int r = r1; r <= r2; ++r)
                for (int c = c1; c <= c2; ++c) grid[r * gridCols + c].push_back(i);
        }
    }

    // 05:144-192: top-most hit of a downward vertical ray.
    double getGroundHeight(double x, double y) const {
        if (triangles.empty()) return -1e30;
        if (x < geoMin.x || x > geoMax.x || y < geoMin.y || y > geoMax.y) return -1e30;
        int r = (int)((x - geoMin.x) / cellW);
        int c = (int)((y - geoMin.y) / cellH);
        if (r < 0 || r >= gridRows || c < 0 || c >= gridCols) return -1e30;
        const auto &candidates = grid[r * gridCols + c];
        if (candidates.empty()) return -1e30;
        const Point ro{x, y, 10000.0};
        const Point rd{0.0, 0.0, -1.0};
        double maxZ = -1e30;
        for (int idx : candidates) {
            const auto &tri = triangles[idx];
            if (tri.maxPt.x < x - 1e-6 || tri.minPt.x > x + 1e-6) continue;
            if (tri.maxPt.y < y - 1e-6 || tri.minPt.y > y + 1e-6) continue;
            const double EPS = 1e-9;
            Point e1{tri.v2.x - tri.v1.x, tri.v2.y - tri.v1.y, tri.v2.z - tri.v1.z};
            Point e2{tri.v3.x - tri.v1.x, tri.v3.y - tri.v1.y, tri.v3.z - tri.v1.z};
            Point h{rd.y * e2.z - rd.z * e2.y, rd.z * e2.x - rd.x * e2.z, rd.x * e2.y - rd.y * e2.x};
            double a = e1.x * h.x + e1.y * h.y + e1.z * h.z;
            if (std::abs(a) < EPS) continue;
            double f = 1.0 / a;
            Point s{ro.x - tri.v1.x, ro.y - tri.v1.y, ro.z - tri.v1.z};
            double u = f * (s.x * h.x + s.y * h.y + s.z * h.z);
            if (u < 0.0 || u > 1.0) continue;
            Point q{s.y * e1.z - s.z * e1.y, s.z * e1.x - s.x * e1.z, s.x * e1.y - s.y * e1.x};
            double v = f * (rd.x * q.x + rd.y * q.y + rd.z * q.z);
            if (v < 0.0 || (u + v) > 1.0) continue;
            double t = f * (e2.x * q.x + e2.y * q.y + e2.z * q.z);
            if (t > EPS) {
                double hitZ = ro.z + t * rd.z;
                if (hitZ > maxZ) maxZ = hitZ;
            }
        }
        return maxZ;
    }
};

inline size_t loc_index(int i, int j, int k, int nx_t, int ny_t) {   // 05:206-209
    return (size_t)i + (size_t)j * (size_t)nx_t + (size_t)k * (size_t)nx_t * (size_t)ny_t;
}

}  // namespace

struct oracle {
    nlos_config cfg{};
    nlos_source src{};
    std::vector<double> samples;
    bool has_source = false;

    int inv_q[Q];
    int NXg = 0, NYg = 0, NZg = 0;
    int nx = 0, ny = 0, nz = 0, nx_t = 0, ny_t = 0, nz_t = 0;
    double dt = 0.0;
    long long step = 0;

    Geometry geo;
    std::vector<double> f_curr, f_next, rho, sponge_sigma;
    std::vector<uint8_t> mask;
    std::vector<double> gH;    // column heights of the last init_domain

    struct Probe { int gi, gj, gk; std::vector<double> p_history; };
    std::vector<Probe> probes;

    std::vector<double> rL, rR, rF, rB, rD, rU;   // receive buffers of the single rank

    // LBMSolver::feq (05:286-291)
    static inline double feq(int k, double r, double ux, double uy, double uz) {
        double cu = ex[k] * ux + ey[k] * uy + ez[k] * uz;
        double u2 = ux * ux + uy * uy + uz * uz;
        return w[k] * r * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * u2);
    }

    bool init(const nlos_config &c) {
        cfg = c;
        // 05:301-312
        for (int i = 0; i < Q; ++i) {
            inv_q[i] = 0;
            for (int j = 0; j < Q; ++j)
                if (ex[i] == -ex[j] && ey[i] == -ey[j] && ez[i] == -ez[j]) { inv_q[i] = j; break; }
        }
        // 05:332-334, single rank: off = 0, n = N (05:354-359)
        NXg = (int)((cfg.x_max - cfg.x_min) / cfg.dx + 0.5);
        NYg = (int)((cfg.y_max - cfg.y_min) / cfg.dx + 0.5);
        NZg = (int)((cfg.z_max - cfg.z_min) / cfg.dx + 0.5);
        if (NXg < 1 || NYg < 1 || NZg < 1) return false;
        nx = NXg; ny = NYg; nz = NZg;
        nx_t = nx + 2; ny_t = ny + 2; nz_t = nz + 2;
        dt = (cs * cfg.dx) / cfg.c_phys;     // UnitConverter::get_dt_phys (05:200-203)
        size_t Nloc = (size_t)nx_t * ny_t * nz_t;
        f_curr.assign(Nloc * Q, 0.0);
        f_next.assign(Nloc * Q, 0.0);
        rho.assign(Nloc, 1.0);
        sponge_sigma.assign(Nloc, 0.0);
        mask.assign(Nloc, FLUID);
        rL.resize((size_t)ny * nz * Q); rR.resize(rL.size());
        rF.resize((size_t)(nx + 2) * nz * Q); rB.resize(rF.size());
        rD.resize((size_t)(nx + 2) * (ny + 2) * Q); rU.resize(rD.size());
        double feq0[Q];
        for (int q = 0; q < Q; ++q) feq0[q] = feq(q, 1.0, 0.0, 0.0, 0.0);
        #pragma omp parallel for
        for (size_t i = 0; i < Nloc; ++i)
            for (int q = 0; q < Q; ++q) f_curr[i * Q + q] = feq0[q];
        init_domain();
        return true;
    }

    // LBMSolver::init_domain (05:453-532); sponge faces/width/strength, sky and ground
    // taken from the configuration.
    void init_domain() {
        const double dxp = cfg.dx;
        const double sponge_w = cfg.sponge_width;
        const double sponge_strength = cfg.sponge_strength;
        const double sky_z = cfg.sky_z;
        const int sf = cfg.sponge_faces;
        gH.assign((size_t)nx_t * ny_t, -1e30);
        #pragma omp parallel for collapse(2)
        for (int j = 0; j < ny_t; ++j)
            for (int i = 0; i < nx_t; ++i) {
                double x = cfg.x_min + ((i - 1) + 0.5) * dxp;
                double y = cfg.y_min + ((j - 1) + 0.5) * dxp;
                gH[(size_t)j * nx_t + i] = geo.getGroundHeight(x, y);
            }
        #pragma omp parallel for collapse(3)
        for (int k = 0; k < nz_t; ++k)
            for (int j = 0; j < ny_t; ++j)
                for (int i = 0; i < nx_t; ++i) {
                    size_t idx = loc_index(i, j, k, nx_t, ny_t);
                    double x = cfg.x_min + ((i - 1) + 0.5) * dxp;
                    double y = cfg.y_min + ((j - 1) + 0.5) * dxp;
                    double z = cfg.z_min + ((k - 1) + 0.5) * dxp;
                    double h = gH[(size_t)j * nx_t + i];
                    if (!cfg.use_geometry) mask[idx] = FLUID;
                    else mask[idx] = (z < h) ? SOLID : FLUID;
                    if (z > sky_z) mask[idx] = FLUID;
                    double d = 1e30;
                    if (sf & NLOS_FACE_XM) d = std::min(d, x - cfg.x_min);
                    if (sf & NLOS_FACE_XP) d = std::min(d, cfg.x_max - x);
                    if (sf & NLOS_FACE_YM) d = std::min(d, y - cfg.y_min);
                    if (sf & NLOS_FACE_YP) d = std::min(d, cfg.y_max - y);
                    if (sf & NLOS_FACE_ZM) d = std::min(d, z - cfg.z_min);
                    if (sf & NLOS_FACE_ZP) d = std::min(d, cfg.z_max - z);
                    if (d < sponge_w) {
                        double r = (sponge_w - d) / sponge_w;
                        sponge_sigma[idx] = sponge_strength * r * r;
                    } else {
                        sponge_sigma[idx] = 0.0;
                    }
                }
        apply_halo_mask();
    }

    // 05:509-531
    void apply_halo_mask() {
        for (int k = 0; k <= nz + 1; ++k)
            for (int j = 0; j <= ny + 1; ++j) {
                mask[loc_index(0, j, k, nx_t, ny_t)] = FLUID;
                mask[loc_index(nx + 1, j, k, nx_t, ny_t)] = FLUID;
            }
        for (int k = 0; k <= nz + 1; ++k)
            for (int i = 0; i <= nx + 1; ++i) {
                mask[loc_index(i, 0, k, nx_t, ny_t)] = FLUID;
                mask[loc_index(i, ny + 1, k, nx_t, ny_t)] = FLUID;
            }
        for (int j = 0; j <= ny + 1; ++j)
            for (int i = 0; i <= nx + 1; ++i) {
                mask[loc_index(i, j, 0, nx_t, ny_t)] = cfg.ground ? SOLID : FLUID;
                mask[loc_index(i, j, nz + 1, nx_t, ny_t)] = FLUID;
            }
    }

    // 05:534-543
    void fill_with_feq0(std::vector<double> &buf) const {
        double feq0[Q];
        for (int q = 0; q < Q; ++q) feq0[q] = feq(q, 1.0, 0.0, 0.0, 0.0);
        size_t cells = buf.size() / Q;
        for (size_t c = 0; c < cells; ++c)
            for (int q = 0; q < Q; ++q) buf[c * Q + q] = feq0[q];
    }

    // 05:545-568
    inline void macro_at(size_t idx, double &r, double &ux, double &uy, double &uz) const {
        r = ux = uy = uz = 0.0;
        const size_t base = idx * Q;
        for (int q = 0; q < Q; ++q) {
            const double fq = f_curr[base + q];
            r += fq;
            ux += fq * ex[q];
            uy += fq * ey[q];
            uz += fq * ez[q];
        }
        if (r > 1e-30) { ux /= r; uy /= r; uz /= r; }
        else { r = 1.0; ux = uy = uz = 0.0; }
    }

    // Source waveform value at t: 05:579-581 (pulse), 08:543 (sine).
    double waveform(double t) const {
        switch (src.waveform) {
        case NLOS_WAVE_GAUSS_PULSE:
            return src.amp * std::exp(-0.5 * std::pow((t - src.t0) / src.sigma_t, 2));
        case NLOS_WAVE_SINE:
            return src.amp * std::sin(2.0 * PI * src.freq * t);
        case NLOS_WAVE_SAMPLED: {
            long long s = std::llround(t / dt);
            if (s < 0 || s >= (long long)samples.size()) return 0.0;
            return src.amp * samples[(size_t)s];
        }
        case NLOS_WAVE_RICKER: {
            double a = PI * src.freq * (t - src.t0);
            a = a * a;
            return src.amp * (1.0 - 2.0 * a) * std::exp(-a);
        }
        }
        return 0.0;
    }

    // DroneSource::update (05:221-238)
    void source_position(double t, double &cx, double &cy, double &cz) const {
        double dx = src.ex - src.sx, dy = src.ey - src.sy, dz = src.ez - src.sz;
        double dist_total = std::sqrt(dx * dx + dy * dy + dz * dz);
        if (dist_total < 1e-12 || src.speed < 1e-12) { cx = src.sx; cy = src.sy; cz = src.sz; return; }
        double travel_time = dist_total / src.speed;
        double r = t / travel_time;
        if (r > 1.0) r = 1.0;
        cx = src.sx + r * dx;
        cy = src.sy + r * dy;
        cz = src.sz + r * dz;
    }

    // LBMSolver::apply_source (05:576-667), one rank (Allreduce of one double is the
    // identity), plus the trilinear variant.
    void apply_source(double t) {
        if (!has_source) return;
        const double DX = cfg.dx;
        const double val = waveform(t);
        const double source_gain = src.gain;
        double cxs, cys, czs;
        source_position(t, cxs, cys, czs);
        if (src.injection == NLOS_INJECT_TRILINEAR) { apply_trilinear(val, cxs, cys, czs); return; }
        const int R = src.stencil_radius;
        const double sigma_cells = src.stencil_sigma_cells;
        const double sigma2 = (sigma_cells * DX) * (sigma_cells * DX);
        const int ic = (int)std::floor((cxs - cfg.x_min) / DX);
        const int jc = (int)std::floor((cys - cfg.y_min) / DX);
        const int kc = (int)std::floor((czs - cfg.z_min) / DX);
        if (ic < -R || ic >= NXg + R || jc < -R || jc >= NYg + R || kc < -R || kc >= NZg + R) return;
        double sumw = 0.0;
        for (int gk = kc - R; gk <= kc + R; ++gk)
            for (int gj = jc - R; gj <= jc + R; ++gj)
                for (int gi = ic - R; gi <= ic + R; ++gi) {
                    if (gi < 0 || gi >= NXg || gj < 0 || gj >= NYg || gk < 0 || gk >= NZg) continue;
                    const double x = cfg.x_min + (gi + 0.5) * DX;
                    const double y = cfg.y_min + (gj + 0.5) * DX;
                    const double z = cfg.z_min + (gk + 0.5) * DX;
                    const double dx = x - cxs, dy = y - cys, dz = z - czs;
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    const double ww = std::exp(-0.5 * r2 / sigma2);
                    const size_t idx = loc_index(gi + 1, gj + 1, gk + 1, nx_t, ny_t);
                    if (mask[idx] == FLUID) sumw += ww;
                }
        if (sumw < 1e-30) return;
        const double A = (val * dt * source_gain) / sumw;
        for (int gk = kc - R; gk <= kc + R; ++gk)
            for (int gj = jc - R; gj <= jc + R; ++gj)
                for (int gi = ic - R; gi <= ic + R; ++gi) {
                    if (gi < 0 || gi >= NXg || gj < 0 || gj >= NYg || gk < 0 || gk >= NZg) continue;
                    const double x = cfg.x_min + (gi + 0.5) * DX;
                    const double y = cfg.y_min + (gj + 0.5) * DX;
                    const double z = cfg.z_min + (gk + 0.5) * DX;
                    const double dx = x - cxs, dy = y - cys, dz = z - czs;
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    const double ww = std::exp(-0.5 * r2 / sigma2);
                    const size_t idx = loc_index(gi + 1, gj + 1, gk + 1, nx_t, ny_t);
                    if (mask[idx] != FLUID) continue;
                    const double add = A * ww;
                    const size_t base = idx * Q;
                    for (int q = 0; q < Q; ++q) f_curr[base + q] += w[q] * add;
                }
    }

    // Trilinear injection: the eight cells whose centres surround the source point, with
    // weights (1-t) / t per axis, normalised over fluid cells like 05:631.
    void apply_trilinear(double val, double cxs, double cys, double czs) {
        const double DX = cfg.dx;
        const double fx = (cxs - cfg.x_min) / DX - 0.5;
        const double fy = (cys - cfg.y_min) / DX - 0.5;
        const double fz = (czs - cfg.z_min) / DX - 0.5;
        const int i0 = (int)std::floor(fx), j0 = (int)std::floor(fy), k0 = (int)std::floor(fz);
        const double tx = fx - i0, ty = fy - j0, tz = fz - k0;
        double wt[8];
        size_t cell[8];
        bool ok[8];
        double sumw = 0.0;
        for (int c = 0; c < 8; ++c) {
            const int di = c & 1, dj = (c >> 1) & 1, dk = (c >> 2) & 1;
            const int gi = i0 + di, gj = j0 + dj, gk = k0 + dk;
            wt[c] = (di ? tx : 1.0 - tx) * (dj ? ty : 1.0 - ty) * (dk ? tz : 1.0 - tz);
            ok[c] = !(gi < 0 || gi >= NXg || gj < 0 || gj >= NYg || gk < 0 || gk >= NZg);
            cell[c] = ok[c] ? loc_index(gi + 1, gj + 1, gk + 1, nx_t, ny_t) : 0;
            if (ok[c] && mask[cell[c]] != FLUID) ok[c] = false;
            if (ok[c]) sumw += wt[c];
        }
        if (sumw < 1e-30) return;
        const double A = (val * dt * src.gain) / sumw;
        for (int c = 0; c < 8; ++c) {
            if (!ok[c]) continue;
            const double add = A * wt[c];
            for (int q = 0; q < Q; ++q) f_curr[cell[c] * Q + q] += w[q] * add;
        }
    }

    // One open face of 05:684-835.  axis: 0,1,2; plus: outward normal along +axis.
    void nscbc_face(int axis, bool plus, std::vector<double> &buf) {
        auto ghost = [&](size_t idxI, double *out) {
            double rI, uI[3];
            macro_at(idxI, rI, uI[0], uI[1], uI[2]);
            const double pI = cs2 * (rI - 1.0);
            double Wp, Wm;
            if (!plus) { Wp = 0.0; Wm = uI[axis] - pI / cs; }
            else       { Wm = 0.0; Wp = uI[axis] + pI / cs; }
            const double unB = 0.5 * (Wp + Wm);
            const double pB = 0.5 * cs * (Wp - Wm);
            const double rhoB = 1.0 + pB / cs2;
            double uB[3] = {uI[0], uI[1], uI[2]};
            uB[axis] = unB;
            const size_t baseI = idxI * Q;
            for (int q = 0; q < Q; ++q) {
                const double feqI = feq(q, rI, uI[0], uI[1], uI[2]);
                const double feqB = feq(q, rhoB, uB[0], uB[1], uB[2]);
                out[q] = feqB + (f_curr[baseI + q] - feqI);
            }
        };
        size_t cnt = 0;
        if (axis == 0) {          // 05:684-736: j in 1..ny, k in 1..nz
            const int iI = plus ? nx : 1;
            for (int k = 1; k <= nz; ++k)
                for (int j = 1; j <= ny; ++j) { ghost(loc_index(iI, j, k, nx_t, ny_t), &buf[cnt]); cnt += Q; }
        } else if (axis == 1) {   // 05:738-802: i in 0..nx+1 (ends keep w), k in 1..nz
            const int jI = plus ? ny : 1;
            for (int k = 1; k <= nz; ++k)
                for (int i = 0; i <= nx + 1; ++i) {
                    if (i == 0 || i == nx + 1) { cnt += Q; continue; }
                    ghost(loc_index(i, jI, k, nx_t, ny_t), &buf[cnt]); cnt += Q;
                }
        } else {                  // 05:804-835: interior i, j only
            const int kI = plus ? nz : 1;
            for (int j = 0; j <= ny + 1; ++j)
                for (int i = 0; i <= nx + 1; ++i) {
                    if (i == 0 || i == nx + 1 || j == 0 || j == ny + 1) { cnt += Q; continue; }
                    ghost(loc_index(i, j, kI, nx_t, ny_t), &buf[cnt]); cnt += Q;
                }
        }
    }

    // LBMSolver::exchange_halos (05:671-904) at one rank: every receive is from
    // MPI_PROC_NULL, so each halo layer ends up holding its receive buffer.
    void exchange_halos() {
        fill_with_feq0(rL); fill_with_feq0(rR); fill_with_feq0(rF);
        fill_with_feq0(rB); fill_with_feq0(rD); fill_with_feq0(rU);
        const int nf = cfg.nscbc_faces;
        if (nf & NLOS_FACE_XM) nscbc_face(0, false, rL);
        if (nf & NLOS_FACE_XP) nscbc_face(0, true, rR);
        if (nf & NLOS_FACE_YM) nscbc_face(1, false, rF);
        if (nf & NLOS_FACE_YP) nscbc_face(1, true, rB);
        if (nf & NLOS_FACE_ZM) nscbc_face(2, false, rD);
        if (nf & NLOS_FACE_ZP) nscbc_face(2, true, rU);
        size_t cnt = 0;
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int q = 0; q < Q; ++q) f_curr[loc_index(0, j, k, nx_t, ny_t) * Q + q] = rL[cnt++];
        cnt = 0;
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int q = 0; q < Q; ++q) f_curr[loc_index(nx + 1, j, k, nx_t, ny_t) * Q + q] = rR[cnt++];
        cnt = 0;
        for (int k = 1; k <= nz; ++k)
            for (int i = 0; i <= nx + 1; ++i)
                for (int q = 0; q < Q; ++q) f_curr[loc_index(i, 0, k, nx_t, ny_t) * Q + q] = rF[cnt++];
        cnt = 0;
        for (int k = 1; k <= nz; ++k)
            for (int i = 0; i <= nx + 1; ++i)
                for (int q = 0; q < Q; ++q) f_curr[loc_index(i, ny + 1, k, nx_t, ny_t) * Q + q] = rB[cnt++];
        cnt = 0;
        for (int j = 0; j <= ny + 1; ++j)
            for (int i = 0; i <= nx + 1; ++i)
                for (int q = 0; q < Q; ++q) f_curr[loc_index(i, j, 0, nx_t, ny_t) * Q + q] = rD[cnt++];
        cnt = 0;
        for (int j = 0; j <= ny + 1; ++j)
            for (int i = 0; i <= nx + 1; ++i)
                for (int q = 0; q < Q; ++q) f_curr[loc_index(i, j, nz + 1, nx_t, ny_t) * Q + q] = rU[cnt++];
    }

    // LBMSolver::step_collision_bgk (05:906-972); wall_reflect < 1 scales the deviation
    // of a bounced population from rest (w_q), which is the identity at 1.
    void step_collision_bgk() {
        const double omega = 1.0 / cfg.tau;
        const double alpha = cfg.wall_reflect;
        #pragma omp parallel for collapse(3)
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    size_t dst = loc_index(i, j, k, nx_t, ny_t);
                    double f_in[Q];
                    for (int q = 0; q < Q; ++q) {
                        int ip = i - ex[q], jp = j - ey[q], kp = k - ez[q];
                        size_t srcc = loc_index(ip, jp, kp, nx_t, ny_t);
                        if (mask[srcc] == SOLID) {
                            f_in[q] = f_curr[dst * Q + inv_q[q]];
                            if (alpha != 1.0) f_in[q] = w[q] + alpha * (f_in[q] - w[q]);
                        } else {
                            f_in[q] = f_curr[srcc * Q + q];
                        }
                    }
                    if (mask[dst] == SOLID) {
                        for (int q = 0; q < Q; ++q) f_next[dst * Q + q] = f_in[q];
                        rho[dst] = 1.0;
                        continue;
                    }
                    double r = 0, jx = 0, jy = 0, jz = 0;
                    for (int q = 0; q < Q; ++q) {
                        r += f_in[q];
                        jx += f_in[q] * ex[q];
                        jy += f_in[q] * ey[q];
                        jz += f_in[q] * ez[q];
                    }
                    if (r <= 0.0) r = 1e-12;
                    double ux = jx / r, uy = jy / r, uz = jz / r;
                    rho[dst] = r;
                    double sponge = sponge_sigma[dst];
                    for (int q = 0; q < Q; ++q) {
                        double feq_q = feq(q, r, ux, uy, uz);
                        double f_post = f_in[q] - omega * (f_in[q] - feq_q);
                        if (sponge > 1e-12) {
                            double feq0 = feq(q, 1.0, 0.0, 0.0, 0.0);
                            f_post = (1.0 - sponge) * f_post + sponge * feq0;
                        }
                        f_next[dst * Q + q] = f_post;
                    }
                }
        f_curr.swap(f_next);
    }

    // Loop body of LBMSolver::run (05:1079-1093).
    void run_steps(long long n) {
        for (long long it = 0; it < n; ++it) {
            double t = step * dt;
            apply_source(t);
            exchange_halos();
            step_collision_bgk();
            for (auto &p : probes) {
                size_t idx = loc_index(p.gi + 1, p.gj + 1, p.gk + 1, nx_t, ny_t);
                p.p_history.push_back(rho[idx] - 1.0);
            }
            ++step;
        }
    }
};

extern "C" {

oracle *oracle_create(const nlos_config *cfg) {
    if (!cfg || !(cfg->dx > 0.0) || !(cfg->tau > 0.5)) return nullptr;
    oracle *o = new oracle();
    if (!o->init(*cfg)) { delete o; return nullptr; }
    return o;
}

void oracle_destroy(oracle *o) { delete o; }

void oracle_dims(const oracle *o, int32_t *dims) { dims[0] = o->NXg; dims[1] = o->NYg; dims[2] = o->NZg; }
double oracle_dt(const oracle *o) { return o->dt; }
long long oracle_step_index(const oracle *o) { return o->step; }

// GeometryEngine::load from a buffer, then init_domain.  Returns 0 on success.
int oracle_load_stl(oracle *o, const void *buf, size_t len, double scale, double sx, double sy, double sz) {
    o->geo = Geometry();
    bool ok = o->geo.load((const unsigned char *)buf, len, scale, sx, sy, sz);
    if (!ok) o->geo = Geometry();
    o->init_domain();
    return ok ? 0 : 1;
}

void oracle_ground_heights(const oracle *o, double *out) {
    std::memcpy(out, o->gH.data(), o->gH.size() * sizeof(double));
}

// Interior mask in/out (x fastest).  Setting re-applies the halo rule (05:509-531).
void oracle_set_mask(oracle *o, const uint8_t *m) {
    size_t n = 0;
    for (int k = 1; k <= o->nz; ++k)
        for (int j = 1; j <= o->ny; ++j)
            for (int i = 1; i <= o->nx; ++i) o->mask[loc_index(i, j, k, o->nx_t, o->ny_t)] = m[n++] ? SOLID : FLUID;
    o->apply_halo_mask();
}

void oracle_get_mask(const oracle *o, uint8_t *m) {
    size_t n = 0;
    for (int k = 1; k <= o->nz; ++k)
        for (int j = 1; j <= o->ny; ++j)
            for (int i = 1; i <= o->nx; ++i) m[n++] = o->mask[loc_index(i, j, k, o->nx_t, o->ny_t)];
}

void oracle_get_sponge(const oracle *o, double *s) {
    size_t n = 0;
    for (int k = 1; k <= o->nz; ++k)
        for (int j = 1; j <= o->ny; ++j)
            for (int i = 1; i <= o->nx; ++i) s[n++] = o->sponge_sigma[loc_index(i, j, k, o->nx_t, o->ny_t)];
}

// Initial populations f = feq(rho,0) on fluid cells.
void oracle_set_density(oracle *o, const double *rho) {
    size_t n = 0;
    for (int k = 1; k <= o->nz; ++k)
        for (int j = 1; j <= o->ny; ++j)
            for (int i = 1; i <= o->nx; ++i, ++n) {
                size_t idx = loc_index(i, j, k, o->nx_t, o->ny_t);
                if (o->mask[idx] != FLUID) continue;
                for (int q = 0; q < Q; ++q) o->f_curr[idx * Q + q] = oracle::feq(q, rho[n], 0.0, 0.0, 0.0);
            }
}

void oracle_set_source(oracle *o, const nlos_source *s) {
    o->src = *s;
    o->samples.assign(s->samples ? s->samples : nullptr, s->samples ? s->samples + s->n_samples : nullptr);
    o->src.samples = nullptr;
    o->has_source = true;
}

// Probe cells floor((p-min)/dx) (05:417-419).  Returns the number of probes outside the grid.
int oracle_set_probes(oracle *o, const double *xyz, int32_t n) {
    o->probes.clear();
    int bad = 0;
    for (int p = 0; p < n; ++p) {
        oracle::Probe pr;
        pr.gi = (int)std::floor((xyz[3 * p + 0] - o->cfg.x_min) / o->cfg.dx);
        pr.gj = (int)std::floor((xyz[3 * p + 1] - o->cfg.y_min) / o->cfg.dx);
        pr.gk = (int)std::floor((xyz[3 * p + 2] - o->cfg.z_min) / o->cfg.dx);
        if (pr.gi < 0 || pr.gi >= o->NXg || pr.gj < 0 || pr.gj >= o->NYg || pr.gk < 0 || pr.gk >= o->NZg) { ++bad; continue; }
        o->probes.push_back(pr);
    }
    return bad;
}

void oracle_run(oracle *o, long long n) { o->run_steps(n); }
void oracle_apply_source(oracle *o, double t) { o->apply_source(t); }
void oracle_exchange_halos(oracle *o) { o->exchange_halos(); }
void oracle_step_collision(oracle *o) { o->step_collision_bgk(); }

// Probe histories, probe-major.  Returns steps per probe.
long long oracle_read_probes(const oracle *o, double *out) {
    long long ns = o->probes.empty() ? 0 : (long long)o->probes[0].p_history.size();
    if (out)
        for (size_t p = 0; p < o->probes.size(); ++p)
            std::memcpy(out + p * ns, o->probes[p].p_history.data(), ns * sizeof(double));
    return ns;
}

// float(rho-1) over the interior, x fastest (05:979-985).
void oracle_read_field(const oracle *o, float *out) {
    size_t n = 0;
    for (int k = 1; k <= o->nz; ++k)
        for (int j = 1; j <= o->ny; ++j)
            for (int i = 1; i <= o->nx; ++i) out[n++] = (float)(o->rho[loc_index(i, j, k, o->nx_t, o->ny_t)] - 1.0);
}

// rho (double) over the interior.
void oracle_read_rho(const oracle *o, double *out) {
    size_t n = 0;
    for (int k = 1; k <= o->nz; ++k)
        for (int j = 1; j <= o->ny; ++j)
            for (int i = 1; i <= o->nx; ++i) out[n++] = o->rho[loc_index(i, j, k, o->nx_t, o->ny_t)];
}

// Full f_curr including the halo, AoS [cell][q].
void oracle_read_f(const oracle *o, double *out) {
    std::memcpy(out, o->f_curr.data(), o->f_curr.size() * sizeof(double));
}

void oracle_write_f(oracle *o, const double *in) {
    std::memcpy(o->f_curr.data(), in, o->f_curr.size() * sizeof(double));
}

}  // extern "C"
