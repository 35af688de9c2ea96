// C-ABI implementation of the engine (include/nlos_engine.h): configuration, geometry,
// source schedule, probe and field output, checkpoints, and the per-step launch sequence.
// Host arithmetic that must round like the reference (grid sizes, dt, sponge ramp,
// source weights, STL transform) is written as in 05_sensor_illusion.cpp and built with
// -ffp-contract=off.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/nlos_engine.h"
#include "common.hpp"
#include "launch.hpp"

#define NLOS_API extern "C" __attribute__((visibility("default")))

using namespace nlos;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(NLOS_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                 \
    } while (0)

// Device buffer that grows on demand.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t n) {
        if (n <= bytes) return hipSuccess;
        if (p) { hipError_t e = hipFree(p); if (e != hipSuccess) return e; p = nullptr; bytes = 0; }
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess) bytes = n;
        return e;
    }
    void release() { if (p) (void)hipFree(p); p = nullptr; bytes = 0; }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// A DevBuf local to one call, freed on every return path.
struct ScopedBuf : DevBuf {
    ScopedBuf() = default;
    ScopedBuf(const ScopedBuf &) = delete;
    ScopedBuf &operator=(const ScopedBuf &) = delete;
    ~ScopedBuf() { release(); }
};

struct Triangle13 { double v[13]; };   // v1, v2, v3, min x, min y, max x, max y

// Binary STL parse and 200x200 binning (GeometryEngine::load/buildGrid, 05:70-142).
struct HostGeometry {
    std::vector<Triangle13> tris;
    double gmin[3] = {1e30, 1e30, 1e30}, gmax[3] = {-1e30, -1e30, -1e30};
    int rows = 200, cols = 200;
    double cell_w = 1.0, cell_h = 1.0;
    std::vector<int> bin_start, bin_tris;

    // One 50-byte record (normal, 3 vertices as float32, attribute), transformed as in
    // 05:87-90, with the bounding boxes of 05:92-104.
    void add_record(const unsigned char *r, size_t t, double scale, const double sh[3]) {
        float rec[12];
        std::memcpy(rec, r, 48);
        double v[9];
        for (int p = 0; p < 3; ++p)
            for (int a = 0; a < 3; ++a) v[3 * p + a] = rec[3 + 3 * p + a] * scale + sh[a];
        Triangle13 &T = tris[t];
        for (int a = 0; a < 9; ++a) T.v[a] = v[a];
        double mn[3], mx[3];
        for (int a = 0; a < 3; ++a) {
            mn[a] = std::min({v[a], v[3 + a], v[6 + a]});
            mx[a] = std::max({v[a], v[3 + a], v[6 + a]});
            gmin[a] = std::min(gmin[a], mn[a]);
            gmax[a] = std::max(gmax[a], mx[a]);
        }
        T.v[9] = mn[0]; T.v[10] = mn[1]; T.v[11] = mx[0]; T.v[12] = mx[1];
    }

    bool parse(const unsigned char *buf, size_t len, double scale, double sx, double sy, double sz) {
        tris.clear();
        if (len < 84) return false;
        uint32_t n = 0;
        std::memcpy(&n, buf + 80, 4);
        if (len < 84 + (size_t)n * 50) return false;
        tris.resize(n);
        for (int a = 0; a < 3; ++a) { gmin[a] = 1e30; gmax[a] = -1e30; }
        const double sh[3] = {sx, sy, sz};
        for (uint32_t t = 0; t < n; ++t) add_record(buf + 84 + (size_t)t * 50, t, scale, sh);
        build_bins();
        return true;
    }

    // The same from a file, read in chunks of records (the whole file is never held;
    // GeometryEngine::load reads record by record too, 05:70-111).  False when the file
    // cannot be opened or holds fewer records than its count.
    bool parse_file(const char *path, double scale, double sx, double sy, double sz) {
        tris.clear();
        FILE *f = std::fopen(path, "rb");
        if (!f) return false;
        unsigned char head[84];
        bool ok = std::fread(head, 1, 84, f) == 84;
        uint32_t n = 0;
        if (ok) std::memcpy(&n, head + 80, 4);
        if (ok) {
            tris.resize(n);
            for (int a = 0; a < 3; ++a) { gmin[a] = 1e30; gmax[a] = -1e30; }
            const double sh[3] = {sx, sy, sz};
            constexpr size_t CHUNK = 1 << 16;
            std::vector<unsigned char> buf(CHUNK * 50);
            for (size_t t0 = 0; t0 < n && ok; t0 += CHUNK) {
                const size_t m = std::min<size_t>(CHUNK, n - t0);
                ok = std::fread(buf.data(), 50, m, f) == m;
                for (size_t r = 0; ok && r < m; ++r) add_record(buf.data() + r * 50, t0 + r, scale, sh);
            }
        }
        std::fclose(f);
        if (!ok) { tris.clear(); return false; }
        build_bins();
        return true;
    }

    // CSR bins of the triangles over the plane (b, c) = ((A+1)%3, (A+2)%3) normal to axis A,
    // rows x cols over the mesh's extent there (for the crossing voxeliser).
    struct AxisBins {
        double bmin = 0, bmax = 0, cmin = 0, cmax = 0, cell_b = 1, cell_c = 1;
        std::vector<int> start, list;
    };
    AxisBins axis_bins(int A, int nr = 200, int nc = 200) const {
        AxisBins o;
        const int B = (A + 1) % 3, C = (A + 2) % 3;
        o.bmin = gmin[B]; o.bmax = gmax[B]; o.cmin = gmin[C]; o.cmax = gmax[C];
        o.cell_b = (o.bmax - o.bmin) / nr;
        o.cell_c = (o.cmax - o.cmin) / nc;
        if (o.cell_b < 1e-9) o.cell_b = 1.0;
        if (o.cell_c < 1e-9) o.cell_c = 1.0;
        o.start.assign((size_t)nr * nc + 1, 0);
        auto range = [&](const Triangle13 &T, int &r1, int &r2, int &c1, int &c2) {
            const double b0 = std::min({T.v[B], T.v[3 + B], T.v[6 + B]}), b1 = std::max({T.v[B], T.v[3 + B], T.v[6 + B]});
            const double c0 = std::min({T.v[C], T.v[3 + C], T.v[6 + C]}), c1v = std::max({T.v[C], T.v[3 + C], T.v[6 + C]});
            r1 = std::max(0, std::min(nr - 1, (int)((b0 - o.bmin) / o.cell_b)));
            r2 = std::max(0, std::min(nr - 1, (int)((b1 - o.bmin) / o.cell_b)));
            c1 = std::max(0, std::min(nc - 1, (int)((c0 - o.cmin) / o.cell_c)));
            c2 = std::max(0, std::min(nc - 1, (int)((c1v - o.cmin) / o.cell_c)));
        };
        std::vector<int> count((size_t)nr * nc, 0);
        for (const Triangle13 &T : tris) {
            int r1, r2, c1, c2;
            range(T, r1, r2, c1, c2);
            for (int r = r1; r <= r2; ++r)
                for (int c = c1; c <= c2; ++c) ++count[(size_t)r * nc + c];
        }
        for (size_t b = 0; b < count.size(); ++b) o.start[b + 1] = o.start[b] + count[b];
        o.list.resize(o.start.back());
        std::vector<int> fill(o.start.begin(), o.start.end() - 1);
        for (size_t t = 0; t < tris.size(); ++t) {
            int r1, r2, c1, c2;
            range(tris[t], r1, r2, c1, c2);
            for (int r = r1; r <= r2; ++r)
                for (int c = c1; c <= c2; ++c) o.list[fill[(size_t)r * nc + c]++] = (int)t;
        }
        return o;
    }

    void build_bins() {
        bin_start.assign((size_t)rows * cols + 1, 0);
        bin_tris.clear();
        if (tris.empty()) return;
        cell_w = (gmax[0] - gmin[0]) / rows;
        cell_h = (gmax[1] - gmin[1]) / cols;
        if (cell_w < 1e-9) cell_w = 1.0;
        if (cell_h < 1e-9) cell_h = 1.0;
        std::vector<int> r1(tris.size()), r2(tris.size()), c1(tris.size()), c2(tris.size());
        std::vector<int> count((size_t)rows * cols, 0);
        for (size_t t = 0; t < tris.size(); ++t) {
            const Triangle13 &T = tris[t];
            int a = (int)((T.v[9] - gmin[0]) / cell_w), b = (int)((T.v[11] - gmin[0]) / cell_w);
            int c = (int)((T.v[10] - gmin[1]) / cell_h), d = (int)((T.v[12] - gmin[1]) / cell_h);
            r1[t] = std::max(0, std::min(rows - 1, a));
            r2[t] = std::max(0, std::min(rows - 1, b));
            c1[t] = std::max(0, std::min(cols - 1, c));
            c2[t] = std::max(0, std::min(cols - 1, d));
            for (int r = r1[t]; r <= r2[t]; ++r)
                for (int cc = c1[t]; cc <= c2[t]; ++cc) ++count[(size_t)r * cols + cc];
        }
        for (size_t b = 0; b < count.size(); ++b) bin_start[b + 1] = bin_start[b] + count[b];
        bin_tris.resize(bin_start.back());
        std::vector<int> fill(bin_start.begin(), bin_start.end() - 1);
        for (size_t t = 0; t < tris.size(); ++t)
            for (int r = r1[t]; r <= r2[t]; ++r)
                for (int cc = c1[t]; cc <= c2[t]; ++cc) bin_tris[fill[(size_t)r * cols + cc]++] = (int)t;
    }
};

struct ProbeDesc { double x, y, z; int gi, gj, gk; int unique; bool in_solid; };

}  // namespace

struct nlos_engine {
    nlos_config cfg{};
    int mode = M_F64;
    GridDesc g{};
    double dt = 0.0, omega = 0.0;
    float scale = 1.0f, inv_scale = 1.0f;
    size_t s_bytes = 8, c_bytes = 8;   // storage / compute element sizes
    hipStream_t stream = nullptr;
    int device = 0;

    DevBuf pop, pop2, mask, sig[3], flag, field, hist, probe_cells, s_a, s_b, solid_cells, src_cells,
        src_adds, face_map[6], gh, vox_tri, vox_bins, vox_start, rho_tmp, fb[2], wall_bits;
    DevBuf xh;                        // x-halo records of the blocked two-array modes (StepArgs::xh)
    bool fb_valid = false;           // x-face buffers agree with the lattice
    bool walls_valid = false;        // LAY_P wall copies agree with the lattice
    bool field_valid = false;
    bool field_alloc = false;

    std::vector<uint8_t> hmask;     // interior SOLID flags, x fastest
    std::vector<ProbeDesc> probes;
    std::vector<long long> probe_ucells;
    int n_solid = 0;
    long long hist_cap = 0;         // steps
    long long step = 0;
    int lay = LAY_P;
    int pattern = NLOS_PATTERN_INPLACE;   // resolved stream pattern
    int cur = 0;                          // two-array pattern: which array holds f*

    nlos_source src{};
    std::vector<double> samples;
    bool has_source = false;
    nlos_batch *batch = nullptr;     // set when the engine is a scene of a batch
    std::vector<DevBuf> ck;          // recorded forward: states at rec_start + i * rec_every
    long long rec_start = -1, rec_steps = 0;
    int rec_every = 0;

    dim3 block{64, 4, 1};
    int use_dpp = 1;

    bool timing = false;
    std::vector<hipEvent_t> ev;     // start/stop pairs
    size_t ev_used = 0;
    long long timed_launches = 0;
    double timed_ms = 0.0;

    std::vector<long long> sched_off;
    std::vector<long long> sched_cells;
    std::vector<double> sched_adds;

    // nlos_step of the two-array pattern: every launch of a step finds the step (its source
    // entries, its probe row) through a device counter, so two steps are captured once as a
    // HIP graph and replayed while nothing they read moves (graph_key)
    DevBuf g_ctr;                     // [0] the step, [1] first step of the staged schedule
    DevBuf g_src;                     // SourceArgs of the steps that read pop and pop2
    DevBuf sched_off_d;               // per-step entry offsets of the staged schedule
    long long h_ctr[2] = {0, 0};      // host staging of g_ctr
    SourceArgs h_src[2]{};            // host staging of g_src
    hipGraph_t step_graph = nullptr;
    hipGraphExec_t step_exec = nullptr;
    std::vector<unsigned char> graph_key;   // launch arguments the graph was captured with
    bool step_graph_on = true;        // NLOS_STEP_GRAPH=0: launch every step directly
    // the tracked boundary solids of a step run on side_stream beside its sweep (fork after
    // the ghost layer, join before the next step)
    hipStream_t side_stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;

    // legacy family-A mode (legacy_a.hip)
    int family = NLOS_FAMILY_B;
    DevBuf sigma_a;                   // sponge per interior cell (FP64)
    nlos_line_source line{};          // 04's line source (amp 0: none)
    DevBuf line_cells, line_spatial;  // its fluid cells and exp(-ln2 d2 / b) per cell
    DevBuf line_factor;               // the same factors per interior cell, x fastest (0 elsewhere)
    DevBuf post_b, fb_b;              // fused family-A step: second post array, bounce sources
    bool post_b_ready = false;        // post_b holds the halo's post-collision values
    bool fuse_a = true;               // NLOS_FUSED_A=0: two passes per family-A step
    long long n_line = 0;
    DevBuf field_a;                   // rho per interior cell of the last capture step (FP64)
    std::vector<double> rho0;         // density set before the first step (t = 0 probe line)

    // field dump off the step stream (nlos_dump_field_async)
    hipStream_t copy_stream = nullptr;
    hipEvent_t dump_ready = nullptr, dump_copied = nullptr;
    float *dump_host = nullptr;       // pinned, n_interior floats
    std::thread dump_thread;
    std::string dump_path, dump_msg;
    int dump_format = 0, dump_rc = 0;
    bool dump_pending = false;        // a capture must wait on the device for dump_copied

    long long n_interior() const { return (long long)g.nx * g.ny * g.nz; }
    void *curp() const { return cur == 0 ? pop.p : pop2.p; }
    void *nxtp() const { return cur == 0 ? pop2.p : pop.p; }
    long long cell(int i, int j, int k) const {
        return cell_index(g, i, j, k);
    }
    bool solid_interior(int gi, int gj, int gk) const {
        return hmask[(size_t)gi + (size_t)g.nx * ((size_t)gj + (size_t)g.ny * gk)] != 0;
    }
};

namespace {

// Bytes of one population array: 19 slots per cell; B16 keeps x_0 in FP32, 18 16-bit
// mantissas and an exponent byte (41 bytes per cell, see Mode in common.hpp).
size_t pop_bytes(const nlos_engine *e) {
    if (e->mode == M_B16) return (size_t)e->g.ncell * 41;
    return (size_t)e->g.ncell * Q * e->s_bytes;
}

// ---- dispatch over storage modes ------------------------------------------------------
template <class F>
hipError_t by_mode(int mode, F &&f) {
    switch (mode) {
    case M_F64: return f(Launch<M_F64>{});
    case M_F32: return f(Launch<M_F32>{});
    case M_B16: return f(Launch<M_B16>{});
    default: return f(Launch<M_F16>{});
    }
}

// Per-axis sponge ramp, evaluated as in 05:491-506: sigma(min d) = max over axes of
// sigma(d_axis) because the ramp is monotone in d and rounding is monotone.
std::vector<double> sponge_axis(const nlos_config &c, int n, double lo, double hi, int bit_lo, int bit_hi) {
    std::vector<double> s((size_t)n + 2, 0.0);
    for (int i = 0; i < n + 2; ++i) {
        const double x = lo + ((i - 1) + 0.5) * c.dx;
        double d = 1e30;
        if (c.sponge_faces & bit_lo) d = std::min(d, x - lo);
        if (c.sponge_faces & bit_hi) d = std::min(d, hi - x);
        if (d < c.sponge_width) {
            const double r = (c.sponge_width - d) / c.sponge_width;
            s[i] = c.sponge_strength * r * r;
        }
    }
    return s;
}

int upload_sponge(nlos_engine *e) {
    const nlos_config &c = e->cfg;
    std::vector<double> ax[3] = {
        sponge_axis(c, e->g.nx, c.x_min, c.x_max, NLOS_FACE_XM, NLOS_FACE_XP),
        sponge_axis(c, e->g.ny, c.y_min, c.y_max, NLOS_FACE_YM, NLOS_FACE_YP),
        sponge_axis(c, e->g.nz, c.z_min, c.z_max, NLOS_FACE_ZM, NLOS_FACE_ZP)};
    for (int a = 0; a < 3; ++a) {
        const size_t n = ax[a].size();
        HIP_TRY(e->sig[a].reserve(n * 8));
        if (e->mode == M_F64) {
            HIP_TRY(hipMemcpy(e->sig[a].p, ax[a].data(), n * 8, hipMemcpyHostToDevice));
        } else {
            std::vector<float> f(n);
            for (size_t i = 0; i < n; ++i) f[i] = ax[a][i] > 1e-12 ? (float)ax[a][i] : 0.0f;
            HIP_TRY(hipMemcpy(e->sig[a].p, f.data(), n * 4, hipMemcpyHostToDevice));
        }
    }
    return NLOS_OK;
}

// Family A's default sponge (01:207-225): on the x and y faces, d = how far the cell
// centre lies inside the layer of thickness T, sigma = S*(max(dx, dy)/T)^2 when positive.
int upload_sponge_a(nlos_engine *e) {
    const nlos_config &c = e->cfg;
    const double T = c.sponge_width, S = c.sponge_strength;
    std::vector<double> sg((size_t)e->n_interior(), 0.0);
    size_t n = 0;
    for (int k = 1; k <= e->g.nz; ++k)
        for (int j = 1; j <= e->g.ny; ++j)
            for (int i = 1; i <= e->g.nx; ++i, ++n) {
                const double x = c.x_min + ((i - 1) + 0.5) * c.dx;
                const double y = c.y_min + ((j - 1) + 0.5) * c.dx;
                double dxs = 0.0, dys = 0.0;
                if (x < c.x_min + T) dxs = (c.x_min + T) - x;
                else if (x > c.x_max - T) dxs = x - (c.x_max - T);
                if (y < c.y_min + T) dys = (c.y_min + T) - y;
                else if (y > c.y_max - T) dys = y - (c.y_max - T);
                const double dm = std::max(dxs, dys);
                if (dm > 0.0) {
                    const double r = dm / T;
                    sg[n] = S * r * r;
                }
            }
    HIP_TRY(e->sigma_a.reserve(sg.size() * 8));
    HIP_TRY(hipMemcpy(e->sigma_a.p, sg.data(), sg.size() * 8, hipMemcpyHostToDevice));
    return NLOS_OK;
}

// Probe bits, NEAR bits, tracked solid cells on open faces; state reset to rest.
int finalize_mask(nlos_engine *e) {
    GridDesc &g = e->g;
    e->fb_valid = false;
    e->walls_valid = false;
    if (e->family == NLOS_FAMILY_A) HIP_TRY(launch_mask_periodic_z(e->mask.as<uint8_t>(), g, e->stream));
    HIP_TRY(launch_mask_flags(e->mask.as<uint8_t>(), g, e->stream));
    if (e->wall_bits.p) HIP_TRY(launch_wall_bits(e->mask.as<uint8_t>(), g, e->wall_bits.as<uint32_t>(), e->stream));
    for (long long c : e->probe_ucells) {
        uint8_t b = 0;
        HIP_TRY(hipMemcpyAsync(&b, e->mask.as<uint8_t>() + c, 1, hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
        b |= MK_PROBE;
        HIP_TRY(hipMemcpyAsync(e->mask.as<uint8_t>() + c, &b, 1, hipMemcpyHostToDevice, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    }
    // Solid interior cells on the boundary plane of an NSCBC face: their pulled state
    // enters the ghost through macro_at (05:689-707), so it is tracked.
    std::unordered_map<long long, int> slot_of;
    std::vector<long long> cells;
    std::vector<int> maps[6];
    const int nf = e->cfg.nscbc_faces;
    for (int face = 0; face < 6; ++face) {
        if (!(nf & (1 << face))) continue;
        const int axis = face / 2;
        const bool plus = face & 1;
        const int na = axis == 0 ? g.ny : g.nx;
        const int nb = axis == 2 ? g.ny : g.nz;
        maps[face].assign((size_t)na * nb, -1);
        for (int b = 0; b < nb; ++b)
            for (int a = 0; a < na; ++a) {
                int gi, gj, gk;
                if (axis == 0) { gi = plus ? g.nx - 1 : 0; gj = a; gk = b; }
                else if (axis == 1) { gi = a; gj = plus ? g.ny - 1 : 0; gk = b; }
                else { gi = a; gj = b; gk = plus ? g.nz - 1 : 0; }
                if (!e->solid_interior(gi, gj, gk)) continue;
                const long long c = e->cell(gi + 1, gj + 1, gk + 1);
                auto it = slot_of.find(c);
                int s;
                if (it == slot_of.end()) { s = (int)cells.size(); slot_of[c] = s; cells.push_back(c); }
                else s = it->second;
                maps[face][(size_t)a + (size_t)na * b] = s;
            }
    }
    e->n_solid = (int)cells.size();
    for (int face = 0; face < 6; ++face) {
        if (e->n_solid == 0 || maps[face].empty()) continue;
        HIP_TRY(e->face_map[face].reserve(maps[face].size() * 4));
        HIP_TRY(hipMemcpy(e->face_map[face].p, maps[face].data(), maps[face].size() * 4, hipMemcpyHostToDevice));
    }
    if (e->n_solid > 0) {
        const size_t sb = (size_t)e->n_solid * Q * e->c_bytes;
        HIP_TRY(e->s_a.reserve(sb));
        HIP_TRY(e->s_b.reserve(sb));
        HIP_TRY(e->solid_cells.reserve(cells.size() * 8));
        HIP_TRY(hipMemcpy(e->solid_cells.p, cells.data(), cells.size() * 8, hipMemcpyHostToDevice));
        HIP_TRY(by_mode(e->mode, [&](auto L) { return decltype(L)::fill_state_rest(e->s_a.p, e->n_solid, e->stream); }));
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    return NLOS_OK;
}

// Interior SOLID flags of the device mask into e->hmask.
int pull_host_mask(nlos_engine *e) {
    HIP_TRY(hipStreamSynchronize(e->stream));
    std::vector<uint8_t> full((size_t)e->g.ncell);
    HIP_TRY(hipMemcpy(full.data(), e->mask.p, full.size(), hipMemcpyDeviceToHost));
    e->hmask.assign((size_t)e->n_interior(), 0);
    size_t n = 0;
    for (int k = 1; k <= e->g.nz; ++k)
        for (int j = 1; j <= e->g.ny; ++j) {
            const uint8_t *row = full.data() + e->cell(1, j, k);
            for (int i = 0; i < e->g.nx; ++i) e->hmask[n++] = row[i] & MK_SOLID;
        }
    return NLOS_OK;
}

int check_flag(nlos_engine *e) {
    int f = 0;
    HIP_TRY(hipMemcpyAsync(&f, e->flag.p, sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (f & 1) return fail(NLOS_ERR_NONFINITE, "non-finite density produced during stepping");
    return NLOS_OK;
}

int ensure_hist(nlos_engine *e, long long steps) {
    const long long nu = (long long)e->probe_ucells.size();
    if (nu == 0 || steps <= e->hist_cap) return NLOS_OK;
    long long cap = std::max<long long>(steps, std::max<long long>(1024, 2 * e->hist_cap));
    DevBuf nb;
    HIP_TRY(nb.reserve((size_t)cap * nu * 8));
    HIP_TRY(hipMemsetAsync(nb.p, 0, (size_t)cap * nu * 8, e->stream));
    if (e->hist.p && e->hist_cap > 0)
        HIP_TRY(hipMemcpyAsync(nb.p, e->hist.p, (size_t)e->hist_cap * nu * 8, hipMemcpyDeviceToDevice, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->hist.release();
    e->hist = nb;
    e->hist_cap = cap;
    return NLOS_OK;
}

// ---- source schedule (05:576-667, 08:543) ---------------------------------------------
// The waveform divided by amp (the source value is amp times this).
double waveform_unit(const nlos_engine *e, double t) {
    const nlos_source &s = e->src;
    switch (s.waveform) {
    case NLOS_WAVE_GAUSS_PULSE: return std::exp(-0.5 * std::pow((t - s.t0) / s.sigma_t, 2));
    case NLOS_WAVE_SINE: return std::sin(2.0 * PI * s.freq * t);
    case NLOS_WAVE_SAMPLED: {
        const long long k = std::llround(t / e->dt);
        if (k < 0 || k >= (long long)e->samples.size()) return 0.0;
        return e->samples[(size_t)k];
    }
    case NLOS_WAVE_RICKER: {
        double a = PI * s.freq * (t - s.t0);
        a = a * a;
        return (1.0 - 2.0 * a) * std::exp(-a);
    }
    }
    return 0.0;
}

double waveform(const nlos_engine *e, double t) { return e->src.amp * waveform_unit(e, t); }

// DroneSource::update (05:221-238)
void source_at(const nlos_source &s, double t, double &cx, double &cy, double &cz) {
    const double dx = s.ex - s.sx, dy = s.ey - s.sy, dz = s.ez - s.sz;
    const double dist_total = std::sqrt(dx * dx + dy * dy + dz * dz);
    if (dist_total < 1e-12 || s.speed < 1e-12) { cx = s.sx; cy = s.sy; cz = s.sz; return; }
    const double travel_time = dist_total / s.speed;
    double r = t / travel_time;
    if (r > 1.0) r = 1.0;
    cx = s.sx + r * dx;
    cy = s.sy + r * dy;
    cz = s.sz + r * dz;
}

// Derivatives of the adds of one step (for the adjoint): per entry d add / d val (val the
// waveform value) and d add / d c (c the source centre, 3 per entry).
struct SourceDeriv {
    std::vector<double> dval;
    std::vector<double> dc;
};

// The stencil cells and adds of time t into cells_out / adds_out (room for MAX_STENCIL
// entries), and their derivatives when d is set.  Returns the number of entries.
constexpr int MAX_STENCIL = 343;   // (2R+1)^3 at the largest stencil radius, 3
int schedule_cells(nlos_engine *e, double t, long long *cells_out, double *adds_out, SourceDeriv *d = nullptr) {
    if (!e->has_source) return 0;
    const nlos_config &c = e->cfg;
    const nlos_source &s = e->src;
    const double DX = c.dx;
    const double val = waveform(e, t);
    double cxs, cys, czs;
    source_at(s, t, cxs, cys, czs);
    const int NX = e->g.nx, NY = e->g.ny, NZ = e->g.nz;
    long long cells[512];
    double ww[512];
    double gw[512][3];   // d ww / d c
    int n = 0;
    double sumw = 0.0;
    if (s.injection == NLOS_INJECT_TRILINEAR) {
        const double fx = (cxs - c.x_min) / DX - 0.5;
        const double fy = (cys - c.y_min) / DX - 0.5;
        const double fz = (czs - c.z_min) / DX - 0.5;
        const int i0 = (int)std::floor(fx), j0 = (int)std::floor(fy), k0 = (int)std::floor(fz);
        const double tx = fx - i0, ty = fy - j0, tz = fz - k0;
        for (int q = 0; q < 8; ++q) {
            const int di = q & 1, dj = (q >> 1) & 1, dk = (q >> 2) & 1;
            const int gi = i0 + di, gj = j0 + dj, gk = k0 + dk;
            const double wt = (di ? tx : 1.0 - tx) * (dj ? ty : 1.0 - ty) * (dk ? tz : 1.0 - tz);
            if (gi < 0 || gi >= NX || gj < 0 || gj >= NY || gk < 0 || gk >= NZ) continue;
            if (e->solid_interior(gi, gj, gk)) continue;
            sumw += wt;
            cells[n] = e->cell(gi + 1, gj + 1, gk + 1);
            if (d) {
                const double ax = di ? tx : 1.0 - tx, ay = dj ? ty : 1.0 - ty, az = dk ? tz : 1.0 - tz;
                gw[n][0] = (di ? 1.0 : -1.0) / DX * ay * az;
                gw[n][1] = (dj ? 1.0 : -1.0) / DX * ax * az;
                gw[n][2] = (dk ? 1.0 : -1.0) / DX * ax * ay;
            }
            ww[n++] = wt;
        }
    } else {
        const int R = std::min(s.stencil_radius, 3);
        const double sigma2 = (s.stencil_sigma_cells * DX) * (s.stencil_sigma_cells * DX);
        const int ic = (int)std::floor((cxs - c.x_min) / DX);
        const int jc = (int)std::floor((cys - c.y_min) / DX);
        const int kc = (int)std::floor((czs - c.z_min) / DX);
        if (ic < -R || ic >= NX + R || jc < -R || jc >= NY + R || kc < -R || kc >= NZ + R) return 0;
        for (int gk = kc - R; gk <= kc + R; ++gk)
            for (int gj = jc - R; gj <= jc + R; ++gj)
                for (int gi = ic - R; gi <= ic + R; ++gi) {
                    if (gi < 0 || gi >= NX || gj < 0 || gj >= NY || gk < 0 || gk >= NZ) continue;
                    const double x = c.x_min + (gi + 0.5) * DX;
                    const double y = c.y_min + (gj + 0.5) * DX;
                    const double z = c.z_min + (gk + 0.5) * DX;
                    const double dx = x - cxs, dy = y - cys, dz = z - czs;
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    const double w = std::exp(-0.5 * r2 / sigma2);
                    if (e->solid_interior(gi, gj, gk)) continue;
                    sumw += w;
                    cells[n] = e->cell(gi + 1, gj + 1, gk + 1);
                    if (d) {
                        gw[n][0] = w * dx / sigma2;
                        gw[n][1] = w * dy / sigma2;
                        gw[n][2] = w * dz / sigma2;
                    }
                    ww[n++] = w;
                }
    }
    if (sumw < 1e-30) return 0;
    const double A = (val * e->dt * s.gain) / sumw;
    for (int q = 0; q < n; ++q) {
        cells_out[q] = cells[q];
        adds_out[q] = A * ww[q];
    }
    if (d) {
        // add_q = val dt gain ww_q / S, S = sum ww: d/dc = val dt gain (gw_q S - ww_q sum gw) / S^2
        double sg[3] = {0.0, 0.0, 0.0};
        for (int q = 0; q < n; ++q)
            for (int a = 0; a < 3; ++a) sg[a] += gw[q][a];
        const double f = e->dt * s.gain / sumw;
        for (int q = 0; q < n; ++q) {
            d->dval.push_back(f * ww[q]);
            for (int a = 0; a < 3; ++a) d->dc.push_back(val * f * (gw[q][a] - ww[q] * sg[a] / sumw));
        }
    }
    return n;
}

// Appends the stencil cells and adds of time t to the engine's host schedule.
void schedule_step(nlos_engine *e, double t, SourceDeriv *d = nullptr) {
    long long cells[MAX_STENCIL];
    double adds[MAX_STENCIL];
    const int n = schedule_cells(e, t, cells, adds, d);
    e->sched_cells.insert(e->sched_cells.end(), cells, cells + n);
    e->sched_adds.insert(e->sched_adds.end(), adds, adds + n);
}

int upload_schedule(nlos_engine *e) {
    const size_t n = e->sched_cells.size();
    if (n == 0) return NLOS_OK;
    HIP_TRY(e->src_cells.reserve(n * 8));
    HIP_TRY(e->src_adds.reserve(n * 8));
    HIP_TRY(hipMemcpyAsync(e->src_cells.p, e->sched_cells.data(), n * 8, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(e->src_adds.p, e->sched_adds.data(), n * 8, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return NLOS_OK;
}

// ---- launch helpers -------------------------------------------------------------------
int sync_face_buffers(nlos_engine *e);
int build_line_cells(nlos_engine *e);

// Kernel arguments for the state `phase` steps ahead of the current one (0 = current).
// A two-array step swaps the population arrays, and a step with tracked boundary solids
// swaps their two state buffers; phase 1 is the state after one such step.
void *pop_at(const nlos_engine *e, int phase) { return ((e->cur ^ phase) & 1) ? e->pop2.p : e->pop.p; }
void *solid_state_at(const nlos_engine *e, int phase) { return (phase & 1) ? e->s_b.p : e->s_a.p; }
// The ground layer k = 0 is solid halo (05:520-525) unless the ground is open; the ghost
// kernel then has nothing to write there.
int skip_zm(const nlos_engine *e) { return e->cfg.ground && e->family == NLOS_FAMILY_B ? 1 : 0; }
long long halo_cells(const nlos_engine *e) { return ghost_items(e->g, skip_zm(e)); }

SourceArgs source_args(const nlos_engine *e, int phase) {
    SourceArgs a{};
    a.pop = e->pattern == NLOS_PATTERN_TWO_ARRAY ? pop_at(e, phase) : e->pop.p;
    a.fb[0] = e->fb[0].p;
    a.fb[1] = e->fb[1].p;
    a.mask = e->mask.as<uint8_t>();
    a.g = e->g;
    a.alpha = e->cfg.wall_reflect;
    a.scale = e->scale;
    a.inv_scale = e->inv_scale;
    return a;
}

GhostArgs ghost_args(const nlos_engine *e, int phase) {
    GhostArgs a{};
    a.pop = e->pattern == NLOS_PATTERN_TWO_ARRAY ? pop_at(e, phase) : e->pop.p;
    a.fb[0] = e->fb[0].p;
    a.fb[1] = e->fb[1].p;
    a.mask = e->mask.as<uint8_t>();
    a.g = e->g;
    a.nscbc_faces = e->cfg.nscbc_faces;
    a.s_old = solid_state_at(e, phase);
    for (int f = 0; f < 6; ++f) a.face_map[f] = e->face_map[f].as<int>();
    a.scale = e->scale;
    a.inv_scale = e->inv_scale;
    a.n_halo = halo_cells(e);
    a.skip_zm = skip_zm(e);
    a.xh = e->xh.p;
    return a;
}

SolidArgs solid_args(const nlos_engine *e, int phase) {
    SolidArgs s{};
    s.pop = e->pattern == NLOS_PATTERN_TWO_ARRAY ? pop_at(e, phase) : e->pop.p;
    s.mask = e->mask.as<uint8_t>();
    s.g = e->g;
    s.lay = e->lay;
    s.cells = e->solid_cells.as<long long>();
    s.n = e->n_solid;
    s.s_old = solid_state_at(e, phase);
    s.s_new = solid_state_at(e, phase + 1);
    s.alpha = e->cfg.wall_reflect;
    s.scale = e->scale;
    s.inv_scale = e->inv_scale;
    s.xh = e->xh.p;
    return s;
}

StepArgs step_args(const nlos_engine *e, int phase, dim3 block) {
    StepArgs a{};
    const bool two = e->pattern == NLOS_PATTERN_TWO_ARRAY;
    a.pop = two ? pop_at(e, phase) : e->pop.p;
    a.pop_out = two ? pop_at(e, phase + 1) : nullptr;
    a.mask = e->mask.as<uint8_t>();
    a.sig_x = e->sig[0].p;
    a.sig_y = e->sig[1].p;
    a.sig_z = e->sig[2].p;
    a.g = e->g;
    a.omega = e->omega;
    a.alpha = e->cfg.wall_reflect;
    a.scale = e->scale;
    a.inv_scale = e->inv_scale;
    a.probe_cells = e->probe_cells.as<long long>();
    a.n_probe_cells = (int)e->probe_ucells.size();
    a.flag = e->flag.as<int>();
    a.tiles_x = (e->g.nx + (int)block.x - 1) / (int)block.x;
    a.tiles_y = (e->g.ny + (int)block.y - 1) / (int)block.y;
    a.nblocks = a.tiles_x * a.tiles_y * e->g.nz;
    // lane shifts need full waves: rows of exactly 64-cell tiles, no partial block rows
    a.dpp = e->use_dpp && block.x == 64 && e->g.nx % 64 == 0 && e->g.ny % (int)block.y == 0;
    a.fb[0] = e->fb[0].p;
    a.fb[1] = e->fb[1].p;
    a.wall_bits = e->wall_bits.as<uint32_t>();
    a.xh = e->xh.p;
    return a;
}

// Source adds of entries off .. off+n-1 of the schedule on the device: the engine's schedule
// buffers, or the arrays given.
int launch_source(nlos_engine *e, long long off, long long n, const long long *cells = nullptr,
                  const double *adds = nullptr) {
    if (n <= 0) return NLOS_OK;
    if (int rc = sync_face_buffers(e)) return rc;
    SourceArgs a = source_args(e, 0);
    a.cells = (cells ? cells : e->src_cells.as<long long>()) + off;
    a.adds = (adds ? adds : e->src_adds.as<double>()) + off;
    a.n = (int)n;
    HIP_TRY(by_mode(e->mode, [&](auto L) { return decltype(L)::source(a, e->lay, e->stream); }));
    return NLOS_OK;
}

int sync_face_buffers(nlos_engine *e) {
    if (e->fb_valid || (!e->fb[0].p && !e->fb[1].p)) { e->fb_valid = true; return NLOS_OK; }
    HIP_TRY(by_mode(e->mode, [&](auto L) {
        return decltype(L)::fb_gather(e->curp(), e->mask.as<uint8_t>(), e->g, e->lay, e->fb[0].p, e->fb[1].p,
                                      e->scale, e->inv_scale, e->stream);
    }));
    e->fb_valid = true;
    return NLOS_OK;
}

int launch_halos(nlos_engine *e) {
    if (int rc = sync_face_buffers(e)) return rc;
    const GhostArgs a = ghost_args(e, 0);
    HIP_TRY(by_mode(e->mode, [&](auto L) { return decltype(L)::ghost(a, e->lay, e->stream); }));
    if (e->n_solid > 0) {
        const SolidArgs s = solid_args(e, 0);
        HIP_TRY(by_mode(e->mode, [&](auto L) { return decltype(L)::solid(s, e->stream); }));
        std::swap(e->s_a, e->s_b);
    }
    return NLOS_OK;
}

int launch_sweep_args(nlos_engine *e, const StepArgs &a, bool capture);

// The field buffer of a capture sweep, cleared (solid cells keep 0).
int prepare_capture(nlos_engine *e) {
    // the field buffer may still be on its way to the host for a dump
    if (e->dump_pending) HIP_TRY(hipStreamWaitEvent(e->stream, e->dump_copied, 0));
    if (!e->field_alloc) {
        HIP_TRY(e->field.reserve((size_t)e->n_interior() * 4));
        e->field_alloc = true;
    }
    HIP_TRY(launch_fill_f32(e->field.as<float>(), e->n_interior(), 0.0f, e->stream));
    return NLOS_OK;
}

int launch_sweep(nlos_engine *e, bool capture) {
    if (capture)
        if (int rc = prepare_capture(e)) return rc;
    StepArgs a = step_args(e, 0, e->block);
    a.probe_row = a.n_probe_cells ? e->hist.as<double>() + (size_t)e->step * a.n_probe_cells : nullptr;
    a.field = capture ? e->field.as<float>() : nullptr;
    const bool two = e->pattern == NLOS_PATTERN_TWO_ARRAY;
    const bool odd = e->lay == LAY_P;
    if (!two && odd && !e->walls_valid)
        HIP_TRY(by_mode(e->mode, [&](auto L) {
            return decltype(L)::prime_walls(e->curp(), e->mask.as<uint8_t>(), e->g, e->lay, e->cfg.wall_reflect,
                                            e->stream);
        }));
    e->walls_valid = true;
    if (int rc = launch_sweep_args(e, a, capture)) return rc;
    if (two) e->cur ^= 1;
    else e->lay = odd ? LAY_Q : LAY_P;
    if (capture) e->field_valid = true;
    ++e->step;
    return NLOS_OK;
}

// The sweep kernel with its timing events, if on; no change of host state.
int launch_sweep_args(nlos_engine *e, const StepArgs &a, bool capture) {
    const bool two = e->pattern == NLOS_PATTERN_TWO_ARRAY;
    const bool odd = e->lay == LAY_P;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (e->timing) {
        if (e->ev_used + 2 > e->ev.size()) {
            for (int q = 0; q < 64; ++q) {
                hipEvent_t x;
                HIP_TRY(hipEventCreate(&x));
                e->ev.push_back(x);
            }
        }
        t0 = e->ev[e->ev_used];
        t1 = e->ev[e->ev_used + 1];
        e->ev_used += 2;
        HIP_TRY(hipEventRecord(t0, e->stream));
    }
    if (two) {
        HIP_TRY(by_mode(e->mode, [&](auto L) { return decltype(L)::step_ab(a, capture, e->block, e->stream); }));
    } else {
        HIP_TRY(by_mode(e->mode, [&](auto L) { return decltype(L)::step(a, odd, capture, e->block, e->stream); }));
    }
    if (e->timing) HIP_TRY(hipEventRecord(t1, e->stream));
    return NLOS_OK;
}

// One family-A step (legacy_a.hip); the probe row of this step is recorded.
int launch_a(nlos_engine *e, bool capture) {
    if (capture) {
        if (e->dump_pending) HIP_TRY(hipStreamWaitEvent(e->stream, e->dump_copied, 0));
        HIP_TRY(e->field_a.reserve((size_t)e->n_interior() * 8));
    }
    if (e->n_line > 0) {
        // t = s * dt (04:371, 383), the timed loop's s counted from step0
        const double t = (double)(e->step - e->line.step0) * e->line.dt;
        HIP_TRY(launch_a_line(static_cast<double *>(e->pop.p), e->g, e->line_cells.as<long long>(),
                              e->line_spatial.as<double>(), e->n_line, e->line.amp, std::sin(e->line.omega * t),
                              e->line.dt, e->stream));
    }
    HIP_TRY(launch_a_halo(static_cast<double *>(e->pop.p), e->g, e->stream));
    AStepArgs a{};
    a.f = static_cast<double *>(e->pop.p);
    a.post = static_cast<double *>(e->pop2.p);
    a.mask = e->mask.as<uint8_t>();
    a.sigma = e->sigma_a.as<double>();
    a.g = e->g;
    a.tau = e->cfg.tau;
    a.n_probe_cells = (int)e->probe_ucells.size();
    a.probe_cells = a.n_probe_cells ? e->probe_cells.as<long long>() : nullptr;
    a.probe_row = a.n_probe_cells ? e->hist.as<double>() + (size_t)e->step * a.n_probe_cells : nullptr;
    a.field = capture ? e->field_a.as<double>() : nullptr;
    a.flag = e->flag.as<int>();
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (e->timing) {
        if (e->ev_used + 2 > e->ev.size()) {
            for (int q = 0; q < 64; ++q) {
                hipEvent_t x;
                HIP_TRY(hipEventCreate(&x));
                e->ev.push_back(x);
            }
        }
        t0 = e->ev[e->ev_used];
        t1 = e->ev[e->ev_used + 1];
        e->ev_used += 2;
        HIP_TRY(hipEventRecord(t0, e->stream));
    }
    HIP_TRY(launch_a_sweep(a, e->stream));
    if (e->timing) HIP_TRY(hipEventRecord(t1, e->stream));
    if (capture) e->field_valid = true;
    ++e->step;
    return NLOS_OK;
}

// n family-A steps, each one pass of k_a_fused (see AFusedArgs), after a prime that applies
// the first step's line-source add and halo to the state and collides it (the first half of
// launch_a).  Post-collision values alternate between pop2 and post_b; the bounce sources of
// cells next to a wall between pop and fb_b; the last pass writes the new state into pop.
int run_a_fused(nlos_engine *e, long long n, bool capture_last) {
    const GridDesc &g = e->g;
    const size_t bytes = (size_t)g.ncell * Q * 8;
    HIP_TRY(e->post_b.reserve(bytes));
    HIP_TRY(e->fb_b.reserve(bytes));
    double *pop = static_cast<double *>(e->pop.p);
    auto sin_at = [&](long long step) {
        // t = s * dt (04:371, 383), the timed loop's s counted from step0
        return std::sin(e->line.omega * ((double)(step - e->line.step0) * e->line.dt));
    };
    HIP_TRY(launch_a_mark_line(e->mask.as<uint8_t>(), g, e->line_cells.as<long long>(), e->n_line, e->stream));
    if (e->n_line > 0)
        HIP_TRY(launch_a_line(pop, g, e->line_cells.as<long long>(), e->line_spatial.as<double>(), e->n_line,
                              e->line.amp, sin_at(e->step), e->line.dt, e->stream));
    HIP_TRY(launch_a_halo(pop, g, e->stream));
    double *post[2] = {static_cast<double *>(e->pop2.p), e->post_b.as<double>()};
    double *fbs[2] = {pop, e->fb_b.as<double>()};
    HIP_TRY(launch_a_collide(pop, post[0], e->mask.as<uint8_t>(), g, e->cfg.tau, e->stream));
    if (n > 1 && !e->post_b_ready) {
        // the halo cells' post-collision values, which no pass rewrites
        HIP_TRY(launch_a_collide(pop, post[1], e->mask.as<uint8_t>(), g, e->cfg.tau, e->stream));
        e->post_b_ready = true;
    }
    if (capture_last) {
        if (e->dump_pending) HIP_TRY(hipStreamWaitEvent(e->stream, e->dump_copied, 0));
        HIP_TRY(e->field_a.reserve((size_t)e->n_interior() * 8));
    }
    const int nu = (int)e->probe_ucells.size();
    for (long long s = 0; s < n; ++s) {
        const bool last = s == n - 1;
        AFusedArgs a{};
        a.post = post[s & 1];
        a.fb = fbs[s & 1];
        a.f_out = last ? pop : nullptr;
        a.post_out = last ? nullptr : post[(s + 1) & 1];
        a.fb_out = last ? nullptr : fbs[(s + 1) & 1];
        a.mask = e->mask.as<uint8_t>();
        a.wall_bits = e->wall_bits.as<uint32_t>();
        a.sigma = e->sigma_a.as<double>();
        a.line_factor = e->line_factor.as<double>();
        a.amp = e->line.amp;
        a.sin_t = e->n_line > 0 ? sin_at(e->step + 1) : 0.0;
        a.dt = e->line.dt;
        a.g = g;
        a.tau = e->cfg.tau;
        a.n_probe_cells = nu;
        a.probe_cells = nu ? e->probe_cells.as<long long>() : nullptr;
        a.probe_row = nu ? e->hist.as<double>() + (size_t)e->step * nu : nullptr;
        a.field = capture_last && last ? e->field_a.as<double>() : nullptr;
        a.flag = e->flag.as<int>();
        hipEvent_t t0 = nullptr, t1 = nullptr;
        if (e->timing) {
            if (e->ev_used + 2 > e->ev.size()) {
                for (int q = 0; q < 64; ++q) {
                    hipEvent_t x;
                    HIP_TRY(hipEventCreate(&x));
                    e->ev.push_back(x);
                }
            }
            t0 = e->ev[e->ev_used];
            t1 = e->ev[e->ev_used + 1];
            e->ev_used += 2;
            HIP_TRY(hipEventRecord(t0, e->stream));
        }
        HIP_TRY(launch_a_fused(a, e->stream));
        if (e->timing) HIP_TRY(hipEventRecord(t1, e->stream));
        ++e->step;
    }
    if (capture_last) e->field_valid = true;
    return NLOS_OK;
}

// ---- nlos_step of the two-array pattern: counter-indexed launches and the step graph ----
long long stencil_bound(const nlos_engine *e);

SourceMulti step_source(const nlos_engine *e, int phase) {
    SourceMulti m{};
    m.scenes = e->g_src.as<SourceArgs>() + ((e->cur ^ phase) & 1);   // g_src[k]: array k (pop, pop2)
    m.ctr = e->g_ctr.as<long long>();
    m.ctr_base = e->g_ctr.as<long long>() + 1;
    m.step_off = e->sched_off_d.as<long long>();
    m.entry_scene = nullptr;
    m.cells = e->src_cells.as<long long>();
    m.adds = e->src_adds.as<double>();
    return m;
}

// Sweep arguments of the state `phase` steps ahead whose probe row is hist + step * n,
// step read from the device counter.
StepArgs step_args_ctr(const nlos_engine *e, int phase) {
    StepArgs a = step_args(e, phase, e->block);
    a.step_ctr = e->g_ctr.as<long long>();
    a.ctr_lag = 1;                       // the ghost kernel advanced it (launch_step_ctr)
    a.hist = e->hist.as<double>();
    if (!a.hist) a.n_probe_cells = 0;
    return a;
}

// One step `phase` steps ahead of the host state: source entries of the step the device
// counter names, ghost layer (which also advances the counter), tracked boundary solids,
// sweep (probe row: the counter minus one).  Leaves the host state as it is
// (advance_steps moves it).
int launch_step_ctr(nlos_engine *e, int phase, bool capture) {
    const long long me = stencil_bound(e);
    if (me > 0) {
        const SourceMulti m = step_source(e, phase);
        HIP_TRY(by_mode(e->mode, [&](auto L) { return decltype(L)::source_multi(m, me, e->stream); }));
    }
    GhostArgs ga = ghost_args(e, phase);
    ga.bump = e->g_ctr.as<long long>();   // the step counter moves on inside the ghost launch
    HIP_TRY(by_mode(e->mode, [&](auto L) { return decltype(L)::ghost(ga, e->lay, e->stream); }));
    // The tracked solids pull from this step's state and ghosts and write only their own
    // next state, which nothing reads before the next step's ghost layer: they run beside
    // the sweep (at 1024x1024x256 they took 21 us of every step in line).
    const bool side = e->n_solid > 0;
    if (side) {
        const SolidArgs sa = solid_args(e, phase);
        HIP_TRY(hipEventRecord(e->ev_fork, e->stream));
        HIP_TRY(hipStreamWaitEvent(e->side_stream, e->ev_fork, 0));
        HIP_TRY(by_mode(e->mode, [&](auto L) { return decltype(L)::solid(sa, e->side_stream); }));
        HIP_TRY(hipEventRecord(e->ev_join, e->side_stream));
    }
    if (capture)
        if (int rc = prepare_capture(e)) return rc;
    StepArgs a = step_args_ctr(e, phase);
    a.field = capture ? e->field.as<float>() : nullptr;
    if (int rc = launch_sweep_args(e, a, capture)) return rc;
    if (side) HIP_TRY(hipStreamWaitEvent(e->stream, e->ev_join, 0));
    return NLOS_OK;
}

// Host state after n steps of launch_step_ctr: the arrays and the tracked-solid buffers
// swap once per step.
void advance_steps(nlos_engine *e, long long n, bool captured) {
    if (n & 1) {
        e->cur ^= 1;
        if (e->n_solid > 0) std::swap(e->s_a, e->s_b);
    }
    e->step += n;
    if (captured) e->field_valid = true;
}

template <class T>
void key_append(std::vector<unsigned char> &k, const T &v) {
    const unsigned char *p = reinterpret_cast<const unsigned char *>(&v);
    k.insert(k.end(), p, p + sizeof(T));
}

// Everything the two captured steps bake into their launches.
std::vector<unsigned char> step_graph_key(const nlos_engine *e) {
    std::vector<unsigned char> k;
    for (int ph = 0; ph < 2; ++ph) {
        key_append(k, step_args_ctr(e, ph));
        key_append(k, ghost_args(e, ph));
        key_append(k, solid_args(e, ph));
        key_append(k, step_source(e, ph));
    }
    key_append(k, stencil_bound(e));
    key_append(k, e->n_solid);
    key_append(k, e->lay);
    key_append(k, e->block);
    return k;
}

int drop_step_graph(nlos_engine *e) {
    if (e->step_exec) HIP_TRY(hipGraphExecDestroy(e->step_exec));
    if (e->step_graph) HIP_TRY(hipGraphDestroy(e->step_graph));
    e->step_exec = nullptr;
    e->step_graph = nullptr;
    e->graph_key.clear();
    return NLOS_OK;
}

// The graph of two steps (the current state and the one after it), captured again when
// any argument it bakes in has moved.
int ensure_step_graph(nlos_engine *e) {
    std::vector<unsigned char> key = step_graph_key(e);
    if (e->step_exec && key == e->graph_key) return NLOS_OK;
    if (int rc = drop_step_graph(e)) return rc;
    HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_step_ctr(e, 0, false);
    if (!rc) rc = launch_step_ctr(e, 1, false);
    hipGraph_t g = nullptr;
    const hipError_t he = hipStreamEndCapture(e->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (he != hipSuccess) return fail(NLOS_ERR_HIP, "step graph capture: %s", hipGetErrorString(he));
    e->step_graph = g;
    HIP_TRY(hipGraphInstantiate(&e->step_exec, e->step_graph, nullptr, nullptr, 0));
    e->graph_key = std::move(key);
    return NLOS_OK;
}

// Stages the schedule of steps e->step .. e->step + m - 1 (already in e->sched_*) for
// counter-indexed launches: entries, per-step offsets, source arguments, counter.
int stage_counter_schedule(nlos_engine *e) {
    if (int rc = upload_schedule(e)) return rc;
    const size_t nb = e->sched_off.size() * 8;
    if (e->sched_off_d.bytes < nb) {
        e->sched_off_d.release();
        HIP_TRY(e->sched_off_d.reserve(std::max<size_t>(nb, 4097 * 8)));
    }
    if (!e->g_ctr.p) HIP_TRY(e->g_ctr.reserve(2 * sizeof(long long)));
    if (!e->g_src.p) HIP_TRY(e->g_src.reserve(2 * sizeof(SourceArgs)));
    e->h_ctr[0] = e->step;
    e->h_ctr[1] = e->step;
    // per population array (pop, pop2), whichever is current
    e->h_src[0] = source_args(e, e->cur & 1);
    e->h_src[1] = source_args(e, (e->cur ^ 1) & 1);
    HIP_TRY(hipMemcpyAsync(e->sched_off_d.p, e->sched_off.data(), nb, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(e->g_ctr.p, e->h_ctr, sizeof(e->h_ctr), hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(e->g_src.p, e->h_src, sizeof(e->h_src), hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));   // the staging above is host memory of the engine
    return sync_face_buffers(e);
}

// Steps 0 .. m-1 of the staged schedule; the last one captures the field when asked.  Pairs
// of steps replay the step graph unless timing is on (its events sit around every sweep)
// or NLOS_STEP_GRAPH=0; the rest launch directly, with the same kernels and arguments.
int run_counter_steps(nlos_engine *e, long long m, bool capture_last) {
    long long s = 0;
    const long long direct_tail = capture_last ? 1 : 0;
    if (e->step_graph_on && !e->timing && m - direct_tail >= 2) {
        if (int rc = ensure_step_graph(e)) return rc;
        for (; s + 2 <= m - direct_tail; s += 2) {
            HIP_TRY(hipGraphLaunch(e->step_exec, e->stream));
            advance_steps(e, 2, false);
        }
    }
    for (; s < m; ++s) {
        const bool cap = capture_last && s == m - 1;
        if (int rc = launch_step_ctr(e, 0, cap)) return rc;
        advance_steps(e, 1, cap);
    }
    return NLOS_OK;
}

int collect_timing(nlos_engine *e) {
    if (e->ev_used == 0) return NLOS_OK;
    HIP_TRY(hipStreamSynchronize(e->stream));
    for (size_t q = 0; q + 1 < e->ev_used; q += 2) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, e->ev[q], e->ev[q + 1]));
        e->timed_ms += ms;
        ++e->timed_launches;
    }
    e->ev_used = 0;
    return NLOS_OK;
}

int batch_settle(const nlos_batch *b);

// Every entry point that takes an engine starts here.  A scene of a batch is advanced on
// the batch's stream, and nlos_batch_step returns with that work still in flight; the
// scene's own calls (reads, checkpoints, VTK, dumps, debug access) go through the scene's
// stream or synchronous copies, which do not wait for the batch's non-blocking stream.  So
// for a batch scene the batch's stream is drained first.
int ready(const nlos_engine *e) {
    if (!e) return fail(NLOS_ERR_ARG, "null engine");
    if (e->batch) return batch_settle(e->batch);
    return NLOS_OK;
}

// Grid of an nx*ny*nz interior (see GridDesc): rows padded to 64 cells, interior rows
// starting 64-aligned at i = 1, ncell a multiple of 64.
GridDesc make_grid(int nx, int ny, int nz) {
    GridDesc g{};
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.px = ((nx + 2 + 63) / 64) * 64;
    g.py = ny + 2; g.pz = nz + 2;
    g.plane = (long long)g.px * g.py;
    g.origin = 63;
    g.ncell = ((g.origin + g.plane * g.pz + 63) / 64) * 64;
    return g;
}

int parse_block(const char *s, dim3 &b) {
    int x = 0, y = 0;
    if (!s || std::sscanf(s, "%dx%d", &x, &y) != 2) return 0;
    if (x <= 0 || y <= 0 || x * y != 256 || (x % 64) != 0) return 0;
    b = dim3(x, y, 1);
    return 1;
}

}  // namespace

// =========================================================================================
NLOS_API int nlos_abi_version(void) { return NLOS_ABI_VERSION; }
NLOS_API const char *nlos_last_error(void) { return g_err.c_str(); }

NLOS_API int nlos_config_default(nlos_config *c) {
    if (!c) return fail(NLOS_ERR_ARG, "null config");
    std::memset(c, 0, sizeof *c);
    c->x_min = 0.0; c->x_max = 1650.0;           // 05:250-252
    c->y_min = 0.0; c->y_max = 1150.0;
    c->z_min = 0.0; c->z_max = 250.0;
    c->dx = 2.0;                                 // 05:253
    c->c_phys = 340.0;                           // 05:198
    c->tau = 0.55;                               // 05:268
    c->sponge_width = 20.0 * c->dx;              // 05:455
    c->sponge_strength = 0.35;                   // 05:456
    c->sky_z = 120.0;                            // 05:458
    c->wall_reflect = 1.0;
    c->sponge_faces = NLOS_FACES_REFERENCE;
    c->nscbc_faces = NLOS_FACES_REFERENCE;       // use_nscbc = true (05:669)
    c->ground = 1;
    c->use_geometry = 1;                         // 05:248
    c->output_world_coords = 1;                  // 05:255
    c->precision = NLOS_F64;
    return NLOS_OK;
}

NLOS_API int nlos_source_default(nlos_source *s) {
    if (!s) return fail(NLOS_ERR_ARG, "null source");
    std::memset(s, 0, sizeof *s);
    s->t0 = 0.05;                 // 05:579
    s->sigma_t = 0.005;           // 05:580
    s->gain = 1.0;                // 05:583
    s->stencil_radius = 2;        // 05:587
    s->stencil_sigma_cells = 1.25;// 05:588
    s->waveform = NLOS_WAVE_GAUSS_PULSE;
    s->injection = NLOS_INJECT_GAUSS5;
    return NLOS_OK;
}

NLOS_API int nlos_create(const nlos_config *cfg, nlos_engine **out) {
    if (!cfg || !out) return fail(NLOS_ERR_ARG, "null argument");
    *out = nullptr;
    if (!(cfg->dx > 0.0)) return fail(NLOS_ERR_ARG, "dx must be positive");
    if (!(cfg->tau > 0.5)) return fail(NLOS_ERR_ARG, "tau must exceed 0.5 (got %g)", cfg->tau);
    if (!(cfg->c_phys > 0.0)) return fail(NLOS_ERR_ARG, "c_phys must be positive");
    if (cfg->precision < NLOS_F64 || cfg->precision > NLOS_B16) return fail(NLOS_ERR_ARG, "bad precision");
    if (cfg->precision == NLOS_B16 && cfg->stream_pattern == NLOS_PATTERN_INPLACE)
        return fail(NLOS_ERR_UNSUPPORTED, "NLOS_B16 needs the two-array stream pattern");
    if (cfg->family != NLOS_FAMILY_B && cfg->family != NLOS_FAMILY_A) return fail(NLOS_ERR_ARG, "bad family");
    const bool fam_a = cfg->family == NLOS_FAMILY_A;
    if (fam_a && cfg->precision != NLOS_F64) return fail(NLOS_ERR_UNSUPPORTED, "NLOS_FAMILY_A runs in NLOS_F64");
    if (fam_a && cfg->stream_pattern == NLOS_PATTERN_INPLACE)
        return fail(NLOS_ERR_UNSUPPORTED, "NLOS_FAMILY_A needs the two-array stream pattern");
    if (!(cfg->wall_reflect >= 0.0 && cfg->wall_reflect <= 1.0)) return fail(NLOS_ERR_ARG, "wall_reflect outside [0,1]");
    const int NX = (int)((cfg->x_max - cfg->x_min) / cfg->dx + 0.5);   // 05:332-334
    const int NY = (int)((cfg->y_max - cfg->y_min) / cfg->dx + 0.5);
    int NZ = (int)((cfg->z_max - cfg->z_min) / cfg->dx + 0.5);
    if (fam_a && NZ < 3) NZ = 3;                                        // 01:100
    if (NX < 1 || NY < 1 || NZ < 1) return fail(NLOS_ERR_ARG, "empty grid %d x %d x %d", NX, NY, NZ);
    if ((long long)(NY + 2) * (NZ) > 0x7fffffffLL / 64) return fail(NLOS_ERR_ARG, "grid too large");

    nlos_engine *e = new nlos_engine();
    e->cfg = *cfg;
    e->mode = cfg->precision;
    e->family = cfg->family;
    if (fam_a) {
        // family A has no ground halo and no NSCBC; its halos are set by the step itself
        e->cfg.ground = 0;
        e->cfg.nscbc_faces = 0;
        e->cfg.stream_pattern = NLOS_PATTERN_TWO_ARRAY;
    }
    e->g = make_grid(NX, NY, NZ);
    e->dt = (CS * cfg->dx) / cfg->c_phys;     // 05:200-203
    if (fam_a) e->dt = std::sqrt(1.0 / 3.0) * cfg->dx / cfg->c_phys;     // 01:37-41
    e->omega = 1.0 / cfg->tau;                // 05:908
    e->s_bytes = e->mode == M_F64 ? 8 : (e->mode == M_F32 ? 4 : 2);   // per population
    e->c_bytes = e->mode == M_F64 ? 8 : 4;
    if (e->mode == M_F16) {
        const int k = cfg->f16_scale_log2 ? cfg->f16_scale_log2 : 24;
        e->scale = std::ldexp(1.0f, k);
        e->inv_scale = std::ldexp(1.0f, -k);
    }
    const char *block_env = std::getenv("NLOS_BLOCK");
    if (const char *x = std::getenv("NLOS_DPP")) e->use_dpp = std::atoi(x) != 0;

    auto bail = [&](int rc) { nlos_destroy(e); return rc; };
    hipError_t he = hipGetDevice(&e->device);
    if (he != hipSuccess) return bail(fail(NLOS_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(he)));
    he = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (he != hipSuccess) return bail(fail(NLOS_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(he)));
    // the side stream of the tracked boundary solids (launch_step_ctr), on the engine's device
    if (hipStreamCreateWithFlags(&e->side_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&e->ev_join, hipEventDisableTiming) != hipSuccess)
        return bail(fail(NLOS_ERR_HIP, "side stream creation failed"));
    const size_t popb = pop_bytes(e);
    e->pattern = e->cfg.stream_pattern;
    if (e->pattern == NLOS_PATTERN_AUTO) {
        size_t freeb = 0, totalb = 0;
        if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) freeb = 0;
        // two arrays when both fit with room for the mask and work buffers
        const size_t need = 2 * popb + 5 * (size_t)e->g.ncell + ((size_t)4 << 30);
        e->pattern = freeb >= need ? NLOS_PATTERN_TWO_ARRAY : NLOS_PATTERN_INPLACE;
    }
    if (e->pattern != NLOS_PATTERN_INPLACE && e->pattern != NLOS_PATTERN_TWO_ARRAY)
        return bail(fail(NLOS_ERR_ARG, "bad stream_pattern %d", cfg->stream_pattern));
    if (e->mode == M_B16 && e->pattern != NLOS_PATTERN_TWO_ARRAY)
        return bail(fail(NLOS_ERR_NOMEM, "NLOS_B16 needs two population arrays of %.2f GB each", popb / 1e9));
    e->lay = e->pattern == NLOS_PATTERN_TWO_ARRAY ? LAY_AB : LAY_P;
    // Measured on the 1024x1024x256 city: both sweeps run fastest on 256-cell row segments
    // (in place: F16S 5.18 -> 4.74 ms, FP32 10.31 -> 9.50 ms against 64x4 tiles, whose
    // x-shifted stores of the odd sweep half-fill lines that a workgroup on another XCD
    // fills).  The odd sweep's lane shifts need 64x4 tiles (NLOS_BLOCK=64x4); with them it
    // measured the same as plain shifted loads.
    e->block = dim3(256, 1, 1);
    if (block_env) parse_block(block_env, e->block);
    if (e->pop.reserve(popb) != hipSuccess)
        return bail(fail(NLOS_ERR_NOMEM, "cannot allocate %.2f GB of populations", popb / 1e9));
    if (e->pattern == NLOS_PATTERN_TWO_ARRAY && e->pop2.reserve(popb) != hipSuccess)
        return bail(fail(NLOS_ERR_NOMEM, "cannot allocate a second %.2f GB population array", popb / 1e9));
    if (e->pattern == NLOS_PATTERN_TWO_ARRAY && e->wall_bits.reserve((size_t)e->g.ncell * 4) != hipSuccess)
        return bail(fail(NLOS_ERR_NOMEM, "cannot allocate wall bits"));
    // x-halo records; NLOS_XHALO=0 keeps those halo values in the lattice (the same results)
    const char *xh_env = std::getenv("NLOS_XHALO");
    if (e->pattern == NLOS_PATTERN_TWO_ARRAY && !fam_a && !(xh_env && std::atoi(xh_env) == 0)) {
        const size_t xb = (size_t)xh_bytes(e->g, e->mode, e->s_bytes);
        if (e->xh.reserve(xb) != hipSuccess) return bail(fail(NLOS_ERR_NOMEM, "cannot allocate x-halo records"));
        if (hipMemset(e->xh.p, 0, xb) != hipSuccess) return bail(fail(NLOS_ERR_HIP, "memset x-halo records"));
    }
    if (const char *x = std::getenv("NLOS_STEP_GRAPH")) e->step_graph_on = std::atoi(x) != 0;
    if (const char *x = std::getenv("NLOS_FUSED_A")) e->fuse_a = std::atoi(x) != 0;
    if (e->mask.reserve((size_t)e->g.ncell) != hipSuccess || e->flag.reserve(sizeof(int)) != hipSuccess)
        return bail(fail(NLOS_ERR_NOMEM, "cannot allocate mask"));
    for (int f = 0; f < 2; ++f)
        if ((cfg->nscbc_faces >> f) & 1)
            if (e->fb[f].reserve((size_t)NY * NZ * Q * (e->mode == M_B16 ? 4 : e->s_bytes)) != hipSuccess)
                return bail(fail(NLOS_ERR_NOMEM, "cannot allocate x-face buffers"));
    int rc;
    if ((rc = upload_sponge(e)) != NLOS_OK) return bail(rc);
    if (fam_a && (rc = upload_sponge_a(e)) != NLOS_OK) return bail(rc);
    if (hipMemsetAsync(e->flag.p, 0, sizeof(int), e->stream) != hipSuccess)
        return bail(fail(NLOS_ERR_HIP, "memset flag"));
    if (by_mode(e->mode, [&](auto L) { return decltype(L)::fill_rest(e->pop.p, e->g, e->stream); }) != hipSuccess)
        return bail(fail(NLOS_ERR_HIP, "fill_rest launch failed"));
    if (e->pop2.p &&
        by_mode(e->mode, [&](auto L) { return decltype(L)::fill_rest(e->pop2.p, e->g, e->stream); }) != hipSuccess)
        return bail(fail(NLOS_ERR_HIP, "fill_rest launch failed"));
    // Empty geometry: mask from heights -1e30 everywhere (fluid, ground halo solid).
    VoxelArgs v{};
    v.x_min = cfg->x_min; v.y_min = cfg->y_min; v.z_min = cfg->z_min; v.dx = cfg->dx;
    v.sky_z = cfg->sky_z; v.use_geometry = cfg->use_geometry; v.ground = cfg->ground;
    const size_t ncol = (size_t)(NX + 2) * (NY + 2);
    if (e->gh.reserve(ncol * 8) != hipSuccess) return bail(fail(NLOS_ERR_NOMEM, "heights"));
    if (launch_column_heights(v, e->g, e->gh.as<double>(), e->stream) != hipSuccess ||
        launch_mask_from_heights(v, e->g, e->gh.as<double>(), e->mask.as<uint8_t>(), e->stream) != hipSuccess)
        return bail(fail(NLOS_ERR_HIP, "mask init launch failed"));
    if ((rc = pull_host_mask(e)) != NLOS_OK) return bail(rc);
    if ((rc = finalize_mask(e)) != NLOS_OK) return bail(rc);
    *out = e;
    return NLOS_OK;
}

NLOS_API int nlos_destroy(nlos_engine *e) {
    if (!e) return NLOS_OK;
    if (e->batch) return fail(NLOS_ERR_STATE, "scene belongs to a batch; destroy the batch");
    if (e->dump_thread.joinable()) e->dump_thread.join();
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->copy_stream) (void)hipStreamSynchronize(e->copy_stream);
    if (e->dump_host) (void)hipHostFree(e->dump_host);
    if (e->dump_ready) (void)hipEventDestroy(e->dump_ready);
    if (e->dump_copied) (void)hipEventDestroy(e->dump_copied);
    if (e->copy_stream) (void)hipStreamDestroy(e->copy_stream);
    DevBuf *bufs[] = {&e->pop, &e->pop2, &e->mask, &e->sig[0], &e->sig[1], &e->sig[2], &e->flag, &e->field,
                      &e->hist, &e->probe_cells, &e->s_a, &e->s_b, &e->solid_cells, &e->src_cells,
                      &e->src_adds, &e->face_map[0], &e->face_map[1], &e->face_map[2],
                      &e->face_map[3], &e->face_map[4], &e->face_map[5], &e->gh, &e->vox_tri,
                      &e->vox_bins, &e->vox_start, &e->rho_tmp, &e->fb[0], &e->fb[1], &e->wall_bits,
                      &e->sigma_a, &e->field_a, &e->xh, &e->line_cells, &e->line_spatial,
                      &e->line_factor, &e->post_b, &e->fb_b};
    for (DevBuf *b : bufs) b->release();
    for (DevBuf &b : e->ck) b.release();
    for (hipEvent_t x : e->ev) (void)hipEventDestroy(x);
    if (e->step_exec) (void)hipGraphExecDestroy(e->step_exec);
    if (e->step_graph) (void)hipGraphDestroy(e->step_graph);
    if (e->side_stream) (void)hipStreamSynchronize(e->side_stream);
    if (e->ev_fork) (void)hipEventDestroy(e->ev_fork);
    if (e->ev_join) (void)hipEventDestroy(e->ev_join);
    if (e->side_stream) (void)hipStreamDestroy(e->side_stream);
    e->g_ctr.release();
    e->g_src.release();
    e->sched_off_d.release();
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return NLOS_OK;
}

NLOS_API int nlos_get_info(const nlos_engine *e, nlos_info *info) {
    if (int rc = ready(e)) return rc;
    if (!info) return fail(NLOS_ERR_ARG, "null info");
    std::memset(info, 0, sizeof *info);
    info->nx = e->g.nx; info->ny = e->g.ny; info->nz = e->g.nz;
    info->precision = e->mode;
    info->dt = e->dt;
    info->stream_pattern = e->pattern;
    info->step = e->step;
    info->n_probes = (int64_t)e->probes.size();
    info->n_solid_boundary = e->n_solid;
    info->bytes_per_update = e->mode == M_B16 ? 2.0 * 41 + 1.0 : 2.0 * Q * (double)e->s_bytes + 1.0;
    size_t tot = 0;
    const DevBuf *bufs[] = {&e->pop, &e->pop2, &e->mask, &e->sig[0], &e->sig[1], &e->sig[2], &e->flag, &e->field,
                            &e->hist, &e->probe_cells, &e->s_a, &e->s_b, &e->solid_cells, &e->src_cells,
                            &e->src_adds, &e->gh, &e->fb[0], &e->fb[1], &e->wall_bits, &e->sigma_a,
                            &e->field_a, &e->xh};
    for (const DevBuf *b : bufs) tot += b->bytes;
    info->device_bytes = tot;
    return NLOS_OK;
}

// ---- geometry --------------------------------------------------------------------------
namespace {

// Mask (and column heights) from a parsed mesh; ok = false means the mesh could not be read
// and empty geometry is used, reported as NLOS_ERR_IO like 05:407-410.
int voxelize_geometry(nlos_engine *e, HostGeometry &geo, bool ok) {
    const int mode = e->cfg.voxelizer;
    const int kind = mode & 3;
    const bool three = (mode & NLOS_VOXEL_3AXIS) != 0;
    if (kind > 2 || (mode & ~7) != 0 || (kind == NLOS_VOXEL_HEIGHT && three))
        return fail(NLOS_ERR_ARG, "bad voxelizer %d", mode);
    VoxelArgs v{};
    v.n_tri = (int)geo.tris.size();
    v.grid_rows = geo.rows; v.grid_cols = geo.cols;
    for (int a = 0; a < 3; ++a) { v.geo_min[a] = geo.gmin[a]; v.geo_max[a] = geo.gmax[a]; }
    v.cell_w = geo.cell_w; v.cell_h = geo.cell_h;
    v.x_min = e->cfg.x_min; v.y_min = e->cfg.y_min; v.z_min = e->cfg.z_min; v.dx = e->cfg.dx;
    v.sky_z = e->cfg.sky_z; v.use_geometry = e->cfg.use_geometry; v.ground = e->cfg.ground;
    if (v.n_tri > 0) {
        HIP_TRY(e->vox_tri.reserve(geo.tris.size() * sizeof(Triangle13)));
        HIP_TRY(e->vox_start.reserve(geo.bin_start.size() * 4));
        HIP_TRY(e->vox_bins.reserve(std::max<size_t>(geo.bin_tris.size(), 1) * 4));
        HIP_TRY(hipMemcpy(e->vox_tri.p, geo.tris.data(), geo.tris.size() * sizeof(Triangle13), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(e->vox_start.p, geo.bin_start.data(), geo.bin_start.size() * 4, hipMemcpyHostToDevice));
        if (!geo.bin_tris.empty())
            HIP_TRY(hipMemcpy(e->vox_bins.p, geo.bin_tris.data(), geo.bin_tris.size() * 4, hipMemcpyHostToDevice));
        v.tri = e->vox_tri.as<double>();
        v.bin_start = e->vox_start.as<int>();
        v.bin_tris = e->vox_bins.as<int>();
    }
    HIP_TRY(launch_column_heights(v, e->g, e->gh.as<double>(), e->stream));
    if (kind == NLOS_VOXEL_HEIGHT) {
        HIP_TRY(launch_mask_from_heights(v, e->g, e->gh.as<double>(), e->mask.as<uint8_t>(), e->stream));
    } else {
        const size_t n = (size_t)e->n_interior();
        ScopedBuf votes, flag, bstart, blist;
        HIP_TRY(votes.reserve(n));
        HIP_TRY(flag.reserve(4));
        HIP_TRY(launch_fill_u8(votes.as<uint8_t>(), (long long)n, 0, e->stream));
        HIP_TRY(hipMemsetAsync(flag.p, 0, 4, e->stream));
        const int axes[3] = {2, 0, 1};
        for (int ai = 0; ai < (three ? 3 : 1); ++ai) {
            const int A = axes[ai];
            CrossArgs c{};
            c.axis = A;
            c.winding = kind == NLOS_VOXEL_WINDING;
            c.lo[0] = e->cfg.x_min; c.lo[1] = e->cfg.y_min; c.lo[2] = e->cfg.z_min;
            c.dx = e->cfg.dx;
            c.n[0] = e->g.nx; c.n[1] = e->g.ny; c.n[2] = e->g.nz;
            // irrational fractions of 1e-7 dx keep rays off lattice-placed edges and vertices
            c.off_b = 1e-7 * e->cfg.dx * 0.6180339887498949;
            c.off_c = 1e-7 * e->cfg.dx * 0.4142135623730950;
            c.votes = votes.as<uint8_t>();
            c.overflow = flag.as<int>();
            HostGeometry::AxisBins ab;
            if (v.n_tri > 0) {
                ab = geo.axis_bins(A);
                HIP_TRY(hipStreamSynchronize(e->stream));   // the previous axis's bins are in use
                HIP_TRY(bstart.reserve(ab.start.size() * 4));
                HIP_TRY(blist.reserve(std::max<size_t>(ab.list.size(), 1) * 4));
                HIP_TRY(hipMemcpy(bstart.p, ab.start.data(), ab.start.size() * 4, hipMemcpyHostToDevice));
                if (!ab.list.empty()) HIP_TRY(hipMemcpy(blist.p, ab.list.data(), ab.list.size() * 4, hipMemcpyHostToDevice));
                c.tri = e->vox_tri.as<double>();
                c.bin_start = bstart.as<int>();
                c.bin_tris = blist.as<int>();
                c.rows = 200; c.cols = 200;
                c.bmin = ab.bmin; c.bmax = ab.bmax; c.cmin = ab.cmin; c.cmax = ab.cmax;
                c.cell_b = ab.cell_b; c.cell_c = ab.cell_c;
            }
            HIP_TRY(launch_crossing_votes(c, e->stream));
        }
        HIP_TRY(launch_mask_from_votes(v, e->g, votes.as<uint8_t>(), three ? 2 : 1, e->mask.as<uint8_t>(), e->stream));
        int over = 0;
        HIP_TRY(hipStreamSynchronize(e->stream));
        HIP_TRY(hipMemcpy(&over, flag.p, 4, hipMemcpyDeviceToHost));
        if (over) return fail(NLOS_ERR_UNSUPPORTED, "a grid line crosses the mesh more than 256 times");
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    int rc = pull_host_mask(e);
    if (rc) return rc;
    rc = finalize_mask(e);
    if (rc) return rc;
    if (!ok) return fail(NLOS_ERR_IO, "STL not loaded, using empty geometry");
    return NLOS_OK;
}

}  // namespace

NLOS_API int nlos_voxelize_stl(nlos_engine *e, const void *stl, size_t len, double scale, double sx,
                               double sy, double sz) {
    if (int rc = ready(e)) return rc;
    if (e->step != 0) return fail(NLOS_ERR_STATE, "geometry can only change before the first step");
    HostGeometry geo;
    const bool ok = stl && geo.parse((const unsigned char *)stl, len, scale, sx, sy, sz);
    if (!ok) geo = HostGeometry();
    return voxelize_geometry(e, geo, ok);
}

NLOS_API int nlos_load_stl(nlos_engine *e, const char *path, double scale, double sx, double sy, double sz) {
    if (int rc = ready(e)) return rc;
    if (e->step != 0) return fail(NLOS_ERR_STATE, "geometry can only change before the first step");
    HostGeometry geo;
    const bool ok = path && geo.parse_file(path, scale, sx, sy, sz);
    if (!ok) geo = HostGeometry();
    const int rc = voxelize_geometry(e, geo, ok);
    if (rc == NLOS_ERR_IO) return fail(NLOS_ERR_IO, "STL not loaded (%s), using empty geometry", path ? path : "null");
    return rc;
}

NLOS_API int nlos_get_ground_heights(nlos_engine *e, double *out) {
    if (int rc = ready(e)) return rc;
    if (!out) return fail(NLOS_ERR_ARG, "null output");
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(hipMemcpy(out, e->gh.p, (size_t)(e->g.nx + 2) * (e->g.ny + 2) * 8, hipMemcpyDeviceToHost));
    return NLOS_OK;
}

NLOS_API int nlos_set_mask(nlos_engine *e, const uint8_t *m) {
    if (int rc = ready(e)) return rc;
    if (!m) return fail(NLOS_ERR_ARG, "null mask");
    if (e->step != 0) return fail(NLOS_ERR_STATE, "geometry can only change before the first step");
    std::vector<uint8_t> full((size_t)e->g.ncell, 0);
    const GridDesc &g = e->g;
    for (int k = 0; k <= g.nz + 1; ++k)
        for (int j = 0; j <= g.ny + 1; ++j)
            for (int i = 0; i <= g.nx + 1; ++i) {
                uint8_t v;
                if (k == 0) v = e->cfg.ground ? MK_SOLID : 0;
                else if (k == g.nz + 1 || i == 0 || i == g.nx + 1 || j == 0 || j == g.ny + 1) v = 0;
                else v = m[(size_t)(i - 1) + (size_t)g.nx * ((size_t)(j - 1) + (size_t)g.ny * (k - 1))] ? MK_SOLID : 0;
                full[(size_t)e->cell(i, j, k)] = v;
            }
    HIP_TRY(hipMemcpy(e->mask.p, full.data(), full.size(), hipMemcpyHostToDevice));
    e->hmask.assign(m, m + e->n_interior());
    for (auto &b : e->hmask) b = b ? 1 : 0;
    if (e->family == NLOS_FAMILY_A && e->line.amp != 0.0)
        if (int rc = build_line_cells(e)) return rc;
    return finalize_mask(e);
}

NLOS_API int nlos_get_mask(nlos_engine *e, uint8_t *m) {
    if (int rc = ready(e)) return rc;
    if (!m) return fail(NLOS_ERR_ARG, "null mask");
    std::memcpy(m, e->hmask.data(), e->hmask.size());
    return NLOS_OK;
}

// ---- initial state, source, probes ----------------------------------------------------------
NLOS_API int nlos_set_sponge(nlos_engine *e, const double *sigma) {
    if (int rc = ready(e)) return rc;
    if (e->family != NLOS_FAMILY_A) return fail(NLOS_ERR_UNSUPPORTED, "per-cell sponges are a NLOS_FAMILY_A feature");
    if (!sigma) return fail(NLOS_ERR_ARG, "null sponge");
    HIP_TRY(hipMemcpy(e->sigma_a.p, sigma, (size_t)e->n_interior() * 8, hipMemcpyHostToDevice));
    return NLOS_OK;
}

namespace {

// The cells and spatial factors of 04's apply_source (04:220-232), in its operation order,
// for the current mask (the reference tests the mask at the add).
int build_line_cells(nlos_engine *e) {
    const nlos_line_source &s = e->line;
    std::vector<long long> cells;
    std::vector<double> spatial;
    std::vector<double> factor;       // per interior cell, for the fused step (k_a_fused)
    if (s.amp != 0.0) {
        factor.assign((size_t)e->n_interior(), 0.0);
        const nlos_config &c = e->cfg;
        const double ln2 = std::log(2.0);
        for (int k = 0; k < e->g.nz; ++k)
            for (int j = 0; j < e->g.ny; ++j)
                for (int i = 0; i < e->g.nx; ++i) {
                    const double x = c.x_min + (double)i * c.dx + 0.5 * c.dx;
                    const double y = c.y_min + (double)j * c.dx + 0.5 * c.dx;
                    const double d2 = (x - s.x0) * (x - s.x0) + (y - s.y0) * (y - s.y0);
                    if (d2 > s.r2_max) continue;
                    if (e->solid_interior(i, j, k)) continue;
                    cells.push_back(e->cell(i + 1, j + 1, k + 1));
                    spatial.push_back(std::exp(-ln2 * d2 / s.b));
                    factor[(size_t)i + (size_t)e->g.nx * ((size_t)j + (size_t)e->g.ny * k)] = spatial.back();
                }
    }
    if (!cells.empty()) {
        HIP_TRY(e->line_cells.reserve(cells.size() * 8));
        HIP_TRY(e->line_spatial.reserve(spatial.size() * 8));
        HIP_TRY(hipMemcpy(e->line_cells.p, cells.data(), cells.size() * 8, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(e->line_spatial.p, spatial.data(), spatial.size() * 8, hipMemcpyHostToDevice));
        HIP_TRY(e->line_factor.reserve(factor.size() * 8));
        HIP_TRY(hipMemcpy(e->line_factor.p, factor.data(), factor.size() * 8, hipMemcpyHostToDevice));
    }
    e->n_line = (long long)cells.size();
    return NLOS_OK;
}

}  // namespace

NLOS_API int nlos_set_line_source(nlos_engine *e, const nlos_line_source *s) {
    if (int rc = ready(e)) return rc;
    if (!s) return fail(NLOS_ERR_ARG, "null line source");
    if (e->family != NLOS_FAMILY_A) return fail(NLOS_ERR_UNSUPPORTED, "the line source is a NLOS_FAMILY_A feature");
    if (s->amp != 0.0 && (!(s->b > 0.0) || !(s->dt > 0.0) || !(s->r2_max >= 0.0)))
        return fail(NLOS_ERR_ARG, "line source needs b > 0, dt > 0 and r2_max >= 0");
    e->line = *s;
    return build_line_cells(e);
}

NLOS_API int nlos_set_density(nlos_engine *e, const double *rho) {
    if (int rc = ready(e)) return rc;
    if (!rho) return fail(NLOS_ERR_ARG, "null density");
    if (e->step != 0) return fail(NLOS_ERR_STATE, "initial density can only be set before the first step");
    const size_t n = (size_t)e->n_interior();
    if (e->family == NLOS_FAMILY_A) e->rho0.assign(rho, rho + n);   // 01's t = 0 probe values
    HIP_TRY(e->rho_tmp.reserve(n * 8));
    HIP_TRY(hipMemcpy(e->rho_tmp.p, rho, n * 8, hipMemcpyHostToDevice));
    HIP_TRY(by_mode(e->mode, [&](auto L) {
        return decltype(L)::set_density(e->curp(), e->mask.as<uint8_t>(), e->g, e->rho_tmp.as<double>(), e->scale, e->stream);
    }));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->rho_tmp.release();
    e->fb_valid = false;
    e->walls_valid = false;
    return NLOS_OK;
}

NLOS_API int nlos_set_source(nlos_engine *e, const nlos_source *s) {
    if (int rc = ready(e)) return rc;
    if (!s) return fail(NLOS_ERR_ARG, "null source");
    if (e->family == NLOS_FAMILY_A) return fail(NLOS_ERR_UNSUPPORTED, "sources are a NLOS_FAMILY_B feature");
    if (s->waveform < NLOS_WAVE_GAUSS_PULSE || s->waveform > NLOS_WAVE_RICKER) return fail(NLOS_ERR_ARG, "bad waveform");
    if (s->injection != NLOS_INJECT_GAUSS5 && s->injection != NLOS_INJECT_TRILINEAR) return fail(NLOS_ERR_ARG, "bad injection");
    if (s->injection == NLOS_INJECT_GAUSS5 && (s->stencil_radius < 0 || s->stencil_radius > 3))
        return fail(NLOS_ERR_ARG, "stencil radius must be in [0,3]");
    if (s->waveform == NLOS_WAVE_SAMPLED && (s->n_samples < 0 || (s->n_samples > 0 && !s->samples)))
        return fail(NLOS_ERR_ARG, "sampled waveform needs samples");
    e->src = *s;
    e->samples.assign(s->samples ? s->samples : nullptr, s->samples ? s->samples + s->n_samples : nullptr);
    e->src.samples = nullptr;
    e->has_source = true;
    return NLOS_OK;
}

NLOS_API int nlos_set_probes(nlos_engine *e, const double *xyz, int32_t n, uint8_t *in_solid) {
    if (int rc = ready(e)) return rc;
    if (n < 0 || (n > 0 && !xyz)) return fail(NLOS_ERR_ARG, "bad probe list");
    if (e->step != 0) return fail(NLOS_ERR_STATE, "probes can only change before the first step");
    std::vector<ProbeDesc> ps;
    for (int p = 0; p < n; ++p) {
        ProbeDesc d{};
        d.x = xyz[3 * p]; d.y = xyz[3 * p + 1]; d.z = xyz[3 * p + 2];
        if (e->family == NLOS_FAMILY_A) {
            d.gi = (int)((d.x - e->cfg.x_min) / e->cfg.dx - 0.5);     // 01:452-454
            d.gj = (int)((d.y - e->cfg.y_min) / e->cfg.dx - 0.5);
            d.gk = (int)((d.z - e->cfg.z_min) / e->cfg.dx - 0.5);
        } else {
            d.gi = (int)std::floor((d.x - e->cfg.x_min) / e->cfg.dx);     // 05:417-419
            d.gj = (int)std::floor((d.y - e->cfg.y_min) / e->cfg.dx);
            d.gk = (int)std::floor((d.z - e->cfg.z_min) / e->cfg.dx);
        }
        if (d.gi < 0 || d.gi >= e->g.nx || d.gj < 0 || d.gj >= e->g.ny || d.gk < 0 || d.gk >= e->g.nz)
            return fail(NLOS_ERR_ARG, "probe %d at (%g,%g,%g) lies outside the grid", p, d.x, d.y, d.z);
        d.in_solid = e->solid_interior(d.gi, d.gj, d.gk);
        ps.push_back(d);
    }
    // Clear old probe bits.
    for (long long c : e->probe_ucells) {
        uint8_t b = 0;
        HIP_TRY(hipMemcpy(&b, e->mask.as<uint8_t>() + c, 1, hipMemcpyDeviceToHost));
        b &= (uint8_t)~MK_PROBE;
        HIP_TRY(hipMemcpy(e->mask.as<uint8_t>() + c, &b, 1, hipMemcpyHostToDevice));
    }
    std::vector<long long> uc;
    for (auto &d : ps) uc.push_back(e->cell(d.gi + 1, d.gj + 1, d.gk + 1));
    std::sort(uc.begin(), uc.end());
    uc.erase(std::unique(uc.begin(), uc.end()), uc.end());
    for (auto &d : ps)
        d.unique = (int)(std::lower_bound(uc.begin(), uc.end(), e->cell(d.gi + 1, d.gj + 1, d.gk + 1)) - uc.begin());
    e->probes = ps;
    e->probe_ucells = uc;
    e->hist.release();
    e->hist_cap = 0;
    if (!uc.empty()) {
        HIP_TRY(e->probe_cells.reserve(uc.size() * 8));
        HIP_TRY(hipMemcpy(e->probe_cells.p, uc.data(), uc.size() * 8, hipMemcpyHostToDevice));
        // a first history, so that steps up to this count allocate nothing (nlos_step_noalloc)
        if (int rc = ensure_hist(e, NLOS_PROBE_HISTORY_RESERVE)) return rc;
    }
    if (in_solid)
        for (int p = 0; p < n; ++p) in_solid[p] = ps[p].in_solid ? 1 : 0;
    return finalize_mask(e);
}

// ---- stepping -----------------------------------------------------------------------------
NLOS_API int nlos_apply_source(nlos_engine *e, double t) {
    if (int rc = ready(e)) return rc;
    if (e->batch) return fail(NLOS_ERR_STATE, "scene belongs to a batch; step it with nlos_batch_step");
    if (e->family == NLOS_FAMILY_A) return fail(NLOS_ERR_UNSUPPORTED, "sources are a NLOS_FAMILY_B feature");
    e->sched_cells.clear();
    e->sched_adds.clear();
    schedule_step(e, t);
    int rc = upload_schedule(e);
    if (rc) return rc;
    return launch_source(e, 0, (long long)e->sched_cells.size());
}

NLOS_API int nlos_exchange_halos(nlos_engine *e) {
    if (int rc = ready(e)) return rc;
    if (e->batch) return fail(NLOS_ERR_STATE, "scene belongs to a batch; step it with nlos_batch_step");
    if (e->family == NLOS_FAMILY_A) {
        HIP_TRY(launch_a_halo(static_cast<double *>(e->pop.p), e->g, e->stream));
        return NLOS_OK;
    }
    return launch_halos(e);
}

NLOS_API int nlos_step_collision(nlos_engine *e) {
    if (int rc = ready(e)) return rc;
    if (e->batch) return fail(NLOS_ERR_STATE, "scene belongs to a batch; step it with nlos_batch_step");
    int rc = ensure_hist(e, e->step + 1);
    if (rc) return rc;
    if (e->family == NLOS_FAMILY_A) {
        // the halo of exchange_halos is refreshed as part of the step: 01 always runs the two
        // together (01:526-527), and doing it twice changes nothing
        return launch_a(e, true);
    }
    return launch_sweep(e, true);
}

NLOS_API int nlos_step(nlos_engine *e, int64_t n, int32_t flags) {
    if (int rc = ready(e)) return rc;
    if (e->batch) return fail(NLOS_ERR_STATE, "scene belongs to a batch; step it with nlos_batch_step");
    if (n < 0) return fail(NLOS_ERR_ARG, "negative step count");
    if (n == 0) return NLOS_OK;
    int rc = ensure_hist(e, e->step + n);
    if (rc) return rc;
    if (e->family == NLOS_FAMILY_A) {
        if (e->fuse_a) {
            if ((rc = run_a_fused(e, n, flags & NLOS_STEP_CAPTURE_FIELD))) return rc;
        } else {
            for (long long s = 0; s < n; ++s)
                if ((rc = launch_a(e, (flags & NLOS_STEP_CAPTURE_FIELD) && s == n - 1))) return rc;
        }
        if (e->timing && (rc = collect_timing(e))) return rc;
        return NLOS_OK;
    }
    const long long chunk = 4096;
    for (long long base = 0; base < n; base += chunk) {
        const long long m = std::min(chunk, n - base);
        e->sched_cells.clear();
        e->sched_adds.clear();
        e->sched_off.assign(1, 0);
        for (long long s = 0; s < m; ++s) {
            schedule_step(e, (double)(e->step + s) * e->dt);      // t = s*dt (05:1079)
            e->sched_off.push_back((long long)e->sched_cells.size());
        }
        if (e->pattern == NLOS_PATTERN_TWO_ARRAY) {
            if ((rc = stage_counter_schedule(e))) return rc;
            if ((rc = run_counter_steps(e, m, (flags & NLOS_STEP_CAPTURE_FIELD) && base + m == n))) return rc;
        } else {
            if ((rc = upload_schedule(e))) return rc;
            for (long long s = 0; s < m; ++s) {
                if ((rc = launch_source(e, e->sched_off[s], e->sched_off[s + 1] - e->sched_off[s]))) return rc;
                if ((rc = launch_halos(e))) return rc;
                const bool cap = (flags & NLOS_STEP_CAPTURE_FIELD) && base + s == n - 1;
                if ((rc = launch_sweep(e, cap))) return rc;
            }
        }
        if (e->timing && (rc = collect_timing(e))) return rc;
    }
    return NLOS_OK;
}

// ---- the step without device allocation (nlos_step_noalloc) --------------------------------
// The source schedule of nlos_step is staged in the engine's own device buffers, which grow
// on demand.  nlos_step_noalloc stages it instead in `scratch`, a file-scope device array:
// every GPU the library runs on holds its own copy until the process ends.  Host code never
// passes the array's name as an address (that is a host-side stand-in); the address of each
// device's copy comes from hipGetSymbolAddress, looked up once per device with that device
// current, and kept in g_scratch.  Layout: NOALLOC_ENTRIES cell indices (int64), then
// NOALLOC_ENTRIES adds (double).  Every byte read is written by the same call first, so
// nothing needs clearing; calls that overlap on one GPU (two engines on two streams) race.
constexpr long long NOALLOC_ENTRIES = NLOS_NOALLOC_SOURCE_ENTRIES;
static __device__ __align__(256) unsigned char scratch[NOALLOC_ENTRIES * 16];

namespace {

constexpr int MAX_DEVICES = 64;
constexpr long long NOALLOC_CHUNK = 4096;        // steps per staged chunk, at most
std::mutex g_scratch_mu;                         // guards g_scratch and the host staging below
unsigned char *g_scratch[MAX_DEVICES] = {};
// host staging of one chunk (static storage: the path allocates no host memory either)
long long g_sched_cells[NOALLOC_ENTRIES];
double g_sched_adds[NOALLOC_ENTRIES];
long long g_sched_off[NOALLOC_CHUNK + 1];

int scratch_address(int device, unsigned char **out) {
    if (device < 0 || device >= MAX_DEVICES) return fail(NLOS_ERR_UNSUPPORTED, "device %d beyond %d", device, MAX_DEVICES);
    if (!g_scratch[device]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(NLOS_ERR_HIP, "hipGetSymbolAddress returned no address for the staging array");
        g_scratch[device] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch[device];
    return NLOS_OK;
}

}  // namespace

NLOS_API int nlos_step_noalloc(nlos_engine *e, int64_t n, int32_t flags) {
    if (int rc = ready(e)) return rc;
    if (e->batch) return fail(NLOS_ERR_STATE, "scene belongs to a batch; step it with nlos_batch_step");
    if (n < 0) return fail(NLOS_ERR_ARG, "negative step count");
    if (n == 0) return NLOS_OK;
    if (e->family != NLOS_FAMILY_B) return fail(NLOS_ERR_UNSUPPORTED, "nlos_step_noalloc runs NLOS_FAMILY_B engines");
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != e->device) return fail(NLOS_ERR_STATE, "the engine's device %d is not current (%d is)", e->device, dev);
    // what nlos_step would allocate on the way
    if (!e->probe_ucells.empty() && e->step + n > e->hist_cap)
        return fail(NLOS_ERR_ARG, "probe history holds %lld steps, %lld requested without allocation "
                    "(nlos_set_probes reserves %d; nlos_step grows it)", e->hist_cap, (long long)(e->step + n),
                    NLOS_PROBE_HISTORY_RESERVE);
    if ((flags & NLOS_STEP_CAPTURE_FIELD) && !e->field_alloc)
        return fail(NLOS_ERR_STATE, "no field buffer yet: capture once with nlos_step before capturing here");
    if (e->timing) return fail(NLOS_ERR_STATE, "timing records allocate events; disable it for nlos_step_noalloc");
    const long long bound = e->has_source ? (e->src.injection == NLOS_INJECT_TRILINEAR ? 8 : MAX_STENCIL) : 1;
    if (bound > NOALLOC_ENTRIES)
        return fail(NLOS_ERR_ARG, "a step of %lld source entries exceeds the staging array of %lld",
                    bound, NOALLOC_ENTRIES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    unsigned char *sp = nullptr;
    if (int rc = scratch_address(dev, &sp)) return rc;
    long long *d_cells = reinterpret_cast<long long *>(sp);
    double *d_adds = reinterpret_cast<double *>(sp + NOALLOC_ENTRIES * 8);
    const long long chunk = std::min(NOALLOC_CHUNK, NOALLOC_ENTRIES / bound);
    for (long long base = 0; base < n; base += chunk) {
        const long long m = std::min(chunk, n - base);
        long long off = 0;
        g_sched_off[0] = 0;
        for (long long s = 0; s < m; ++s) {
            // t = s*dt (05:1079); at most `bound` entries per step, so the chunk fits
            off += schedule_cells(e, (double)(e->step + s) * e->dt, g_sched_cells + off, g_sched_adds + off);
            g_sched_off[s + 1] = off;
        }
        if (off > 0) {
            // the previous chunk's kernels on this stream read the staging array first
            HIP_TRY(hipMemcpyAsync(d_cells, g_sched_cells, off * 8, hipMemcpyHostToDevice, e->stream));
            HIP_TRY(hipMemcpyAsync(d_adds, g_sched_adds, off * 8, hipMemcpyHostToDevice, e->stream));
            HIP_TRY(hipStreamSynchronize(e->stream));
        }
        for (long long s = 0; s < m; ++s) {
            if (int rc = launch_source(e, g_sched_off[s], g_sched_off[s + 1] - g_sched_off[s], d_cells, d_adds))
                return rc;
            if (int rc = launch_halos(e)) return rc;
            const bool cap = (flags & NLOS_STEP_CAPTURE_FIELD) && base + s == n - 1;
            if (int rc = launch_sweep(e, cap)) return rc;
        }
    }
    return NLOS_OK;
}

NLOS_API int nlos_synchronize(nlos_engine *e) {
    if (int rc = ready(e)) return rc;
    HIP_TRY(hipStreamSynchronize(e->stream));
    return check_flag(e);
}

// ---- results ------------------------------------------------------------------------------
NLOS_API int nlos_read_probes(nlos_engine *e, double *out, int64_t max_steps, int64_t *n_steps) {
    if (int rc = ready(e)) return rc;
    const long long ns = std::min<long long>(e->step, max_steps);
    if (n_steps) *n_steps = ns;
    if (int rc = check_flag(e)) return rc;
    const size_t nu = e->probe_ucells.size();
    if (!out || nu == 0 || ns <= 0) return NLOS_OK;
    std::vector<double> h((size_t)ns * nu);
    HIP_TRY(hipMemcpy(h.data(), e->hist.p, h.size() * 8, hipMemcpyDeviceToHost));
    for (size_t p = 0; p < e->probes.size(); ++p)
        for (long long s = 0; s < ns; ++s) out[(size_t)p * max_steps + s] = h[(size_t)s * nu + e->probes[p].unique];
    return NLOS_OK;
}

NLOS_API int nlos_read_field(nlos_engine *e, float *out) {
    if (int rc = ready(e)) return rc;
    if (!out) return fail(NLOS_ERR_ARG, "null output");
    if (!e->field_valid) return fail(NLOS_ERR_STATE, "no field captured (step with NLOS_STEP_CAPTURE_FIELD)");
    if (int rc = check_flag(e)) return rc;
    if (e->family == NLOS_FAMILY_A) {
        std::vector<double> r((size_t)e->n_interior());
        HIP_TRY(hipMemcpy(r.data(), e->field_a.p, r.size() * 8, hipMemcpyDeviceToHost));
        for (size_t q = 0; q < r.size(); ++q) out[q] = (float)(r[q] - 1.0);
        return NLOS_OK;
    }
    HIP_TRY(hipMemcpy(out, e->field.p, (size_t)e->n_interior() * 4, hipMemcpyDeviceToHost));
    return NLOS_OK;
}

NLOS_API int nlos_write_probe_files(nlos_engine *e, const char *dir, const char *const *names) {
    if (int rc = ready(e)) return rc;
    if (!names) return fail(NLOS_ERR_ARG, "null names");
    const long long ns = e->step;
    std::vector<double> buf((size_t)e->probes.size() * std::max<long long>(ns, 1));
    int64_t got = 0;
    if (int rc = nlos_read_probes(e, buf.data(), ns, &got)) return rc;
    for (size_t p = 0; p < e->probes.size(); ++p) {
        std::string path = std::string(dir && *dir ? dir : ".") + "/probe_" + names[p] + ".txt";
        FILE *f = std::fopen(path.c_str(), "w");
        if (!f) return fail(NLOS_ERR_IO, "cannot write %s", path.c_str());
        for (long long s = 0; s < got; ++s) std::fprintf(f, "%.8e\n", buf[(size_t)p * ns + s]);   // 05:1113
        std::fclose(f);
    }
    return NLOS_OK;
}

namespace {

// The VTK file of a captured field v (interior, x fastest).  ASCII follows 05:1045-1053:
// default stream formatting (%g) for the header, "%.6e" per value.  BINARY is the legacy
// VTK binary encoding of the same dataset (big-endian float32 after the header).
int write_vtk_host(const nlos_engine *e, const char *path, const float *v, int format) {
    const long long n = e->n_interior();
    FILE *f = std::fopen(path, format == NLOS_VTK_BINARY ? "wb" : "w");
    if (!f) return fail(NLOS_ERR_IO, "cannot write %s", path);
    std::vector<char> iobuf((size_t)1 << 22);
    std::setvbuf(f, iobuf.data(), _IOFBF, iobuf.size());
    double ox = e->cfg.x_min, oy = e->cfg.y_min, oz = e->cfg.z_min;
    if (e->cfg.output_world_coords) { ox += e->cfg.world_offset[0]; oy += e->cfg.world_offset[1]; oz += e->cfg.world_offset[2]; }
    std::fprintf(f, "# vtk DataFile Version 3.0\nLBM Urban Acoustics V21\n%s\n",
                 format == NLOS_VTK_BINARY ? "BINARY" : "ASCII");
    std::fprintf(f, "DATASET STRUCTURED_POINTS\n");
    std::fprintf(f, "DIMENSIONS %d %d %d\n", e->g.nx, e->g.ny, e->g.nz);
    std::fprintf(f, "ORIGIN %g %g %g\n", ox, oy, oz);
    std::fprintf(f, "SPACING %g %g %g\n", e->cfg.dx, e->cfg.dx, e->cfg.dx);
    std::fprintf(f, "POINT_DATA %zu\n", (size_t)n);
    std::fprintf(f, "SCALARS PressureAmp float 1\nLOOKUP_TABLE default\n");
    bool ok = true;
    if (format == NLOS_VTK_BINARY) {
        std::vector<uint32_t> be((size_t)1 << 16);
        for (long long q0 = 0; q0 < n && ok; q0 += (long long)be.size()) {
            const size_t m = (size_t)std::min<long long>((long long)be.size(), n - q0);
            for (size_t r = 0; r < m; ++r) {
                uint32_t u;
                std::memcpy(&u, v + q0 + (long long)r, 4);
                be[r] = __builtin_bswap32(u);
            }
            ok = std::fwrite(be.data(), 4, m, f) == m;
        }
        ok = ok && std::fputc('\n', f) != EOF;
    } else {
        for (long long q = 0; q < n; ++q) std::fprintf(f, "%.6e\n", (double)v[(size_t)q]);
    }
    ok = !std::ferror(f) && ok;
    ok = std::fclose(f) == 0 && ok;
    return ok ? NLOS_OK : fail(NLOS_ERR_IO, "write failed: %s", path);
}

// 01's frame (01:414-436): rho of every interior cell (1 on solids), written as
// (rho - 1) * 100 with "%.6e", non-finite values as 0.
int write_vtk_a(nlos_engine *e, const char *path) {
    if (!e->field_valid) return fail(NLOS_ERR_STATE, "no field captured (step with NLOS_STEP_CAPTURE_FIELD)");
    if (int rc = check_flag(e)) return rc;
    const size_t n = (size_t)e->n_interior();
    std::vector<double> r(n);
    HIP_TRY(hipMemcpy(r.data(), e->field_a.p, n * 8, hipMemcpyDeviceToHost));
    FILE *f = std::fopen(path, "w");
    if (!f) return fail(NLOS_ERR_IO, "cannot write %s", path);
    std::vector<char> iobuf((size_t)1 << 22);
    std::setvbuf(f, iobuf.data(), _IOFBF, iobuf.size());
    std::fprintf(f, "# vtk DataFile Version 3.0\nLBM Final Corrected\nASCII\nDATASET STRUCTURED_POINTS\n");
    std::fprintf(f, "DIMENSIONS %d %d %d\n", e->g.nx, e->g.ny, e->g.nz);
    std::fprintf(f, "ORIGIN %g %g %g\n", e->cfg.x_min, e->cfg.y_min, e->cfg.z_min);
    std::fprintf(f, "SPACING %g %g %g\n", e->cfg.dx, e->cfg.dx, e->cfg.dx);
    std::fprintf(f, "POINT_DATA %zu\n", n);
    std::fprintf(f, "SCALARS Pressure double 1\nLOOKUP_TABLE default\n");
    for (size_t q = 0; q < n; ++q) {
        double v = r[q];
        if (std::isnan(v) || std::isinf(v)) v = 1.0;
        std::fprintf(f, "%.6e\n", (v - 1.0) * 100.0);
    }
    const bool ok = !std::ferror(f);
    return (std::fclose(f) == 0 && ok) ? NLOS_OK : fail(NLOS_ERR_IO, "write failed: %s", path);
}

int write_vtk_sync(nlos_engine *e, const char *path, int format) {
    if (int rc = ready(e)) return rc;
    if (!path) return fail(NLOS_ERR_ARG, "null path");
    if (e->family == NLOS_FAMILY_A) {
        if (format != NLOS_VTK_ASCII) return fail(NLOS_ERR_UNSUPPORTED, "NLOS_FAMILY_A writes 01's ASCII frames");
        return write_vtk_a(e, path);
    }
    std::vector<float> v((size_t)e->n_interior());
    if (int rc = nlos_read_field(e, v.data())) return rc;
    return write_vtk_host(e, path, v.data(), format);
}

}  // namespace

NLOS_API int nlos_write_vtk(nlos_engine *e, const char *path) { return write_vtk_sync(e, path, NLOS_VTK_ASCII); }

NLOS_API int nlos_write_vtk_binary(nlos_engine *e, const char *path) {
    return write_vtk_sync(e, path, NLOS_VTK_BINARY);
}

NLOS_API int nlos_dump_wait(nlos_engine *e) {
    if (int rc = ready(e)) return rc;
    if (e->dump_thread.joinable()) e->dump_thread.join();
    e->dump_pending = false;
    if (e->dump_rc != NLOS_OK) {
        const int rc = e->dump_rc;
        e->dump_rc = NLOS_OK;
        return fail(rc, "field dump: %s", e->dump_msg.c_str());
    }
    return NLOS_OK;
}

NLOS_API int nlos_dump_field_async(nlos_engine *e, const char *path, int32_t format) {
    if (int rc = ready(e)) return rc;
    if (!path) return fail(NLOS_ERR_ARG, "null path");
    if (format != NLOS_VTK_ASCII && format != NLOS_VTK_BINARY) return fail(NLOS_ERR_ARG, "bad VTK format %d", format);
    if (e->family == NLOS_FAMILY_A) return fail(NLOS_ERR_UNSUPPORTED, "NLOS_FAMILY_A frames are written synchronously");
    if (!e->field_valid) return fail(NLOS_ERR_STATE, "no field captured (step with NLOS_STEP_CAPTURE_FIELD)");
    if (int rc = nlos_dump_wait(e)) return rc;
    const size_t bytes = (size_t)e->n_interior() * 4;
    if (!e->copy_stream) {
        HIP_TRY(hipStreamCreateWithFlags(&e->copy_stream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&e->dump_ready, hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&e->dump_copied, hipEventDisableTiming));
    }
    if (!e->dump_host) HIP_TRY(hipHostMalloc((void **)&e->dump_host, bytes, hipHostMallocDefault));
    HIP_TRY(hipEventRecord(e->dump_ready, e->stream));
    HIP_TRY(hipStreamWaitEvent(e->copy_stream, e->dump_ready, 0));
    HIP_TRY(hipMemcpyAsync(e->dump_host, e->field.p, bytes, hipMemcpyDeviceToHost, e->copy_stream));
    HIP_TRY(hipEventRecord(e->dump_copied, e->copy_stream));
    e->dump_pending = true;
    e->dump_path = path;
    e->dump_format = format;
    e->dump_thread = std::thread([e]() {
        const hipError_t he = hipEventSynchronize(e->dump_copied);
        if (he != hipSuccess) {
            e->dump_rc = NLOS_ERR_HIP;
            e->dump_msg = hipGetErrorString(he);
            return;
        }
        e->dump_rc = write_vtk_host(e, e->dump_path.c_str(), e->dump_host, e->dump_format);
        if (e->dump_rc != NLOS_OK) e->dump_msg = g_err;   // this thread's message
    });
    return NLOS_OK;
}

NLOS_API int nlos_write_probe_dat(nlos_engine *e, const char *dir, const char *const *names, double scale) {
    if (int rc = ready(e)) return rc;
    if (!names) return fail(NLOS_ERR_ARG, "null names");
    const long long ns = e->step;
    std::vector<double> buf((size_t)e->probes.size() * std::max<long long>(ns, 1));
    int64_t got = 0;
    if (int rc = nlos_read_probes(e, buf.data(), ns, &got)) return rc;
    for (size_t p = 0; p < e->probes.size(); ++p) {
        std::string path = std::string(dir && *dir ? dir : ".") + "/probe_" + names[p] + ".dat";
        FILE *f = std::fopen(path.c_str(), "w");
        if (!f) return fail(NLOS_ERR_IO, "cannot write %s", path.c_str());
        std::fprintf(f, "# Time Pressure\n");                                        // 01:487
        if (e->family == NLOS_FAMILY_A) {
            // 01 records before every step, so its first line is the state at t = 0 (01:509)
            const ProbeDesc &d = e->probes[p];
            double v0 = 0.0;
            if (!d.in_solid && !e->rho0.empty())
                v0 = e->rho0[(size_t)d.gi + (size_t)e->g.nx * ((size_t)d.gj + (size_t)e->g.ny * d.gk)] - 1.0;
            std::fprintf(f, "%g %g\n", 0.0, v0 * scale);
        }
        for (long long s = 0; s < got; ++s)                                          // 01:488-491
            std::fprintf(f, "%g %g\n", (double)(s + 1) * e->dt, buf[(size_t)p * ns + s] * scale);
        if (std::fclose(f) != 0) return fail(NLOS_ERR_IO, "write failed: %s", path.c_str());
    }
    return NLOS_OK;
}

// ---- checkpoint / restore -------------------------------------------------------------------
// Format version 3: populations blocked by 64 cells (pop_index); version 2 held them slot by
// slot and is refused.
namespace {
constexpr int32_t CKPT_VERSION = 3;
struct CkptHeader {
    char magic[8];
    int32_t version, mode, family, reserved;
    int32_t nx, ny, nz, lay;
    int64_t step, ncell, n_solid, n_ucells;
    uint64_t pop_bytes, state_bytes, hist_bytes;
};
}

NLOS_API int nlos_checkpoint(nlos_engine *e, void *buf, size_t cap, size_t *needed) {
    if (int rc = ready(e)) return rc;
    CkptHeader h{};
    std::memcpy(h.magic, "NLOSCKP1", 8);
    h.version = CKPT_VERSION; h.mode = e->mode; h.family = e->family;
    h.nx = e->g.nx; h.ny = e->g.ny; h.nz = e->g.nz; h.lay = e->lay;
    h.step = e->step; h.ncell = e->g.ncell; h.n_solid = e->n_solid;
    h.n_ucells = (int64_t)e->probe_ucells.size();
    h.pop_bytes = pop_bytes(e);
    h.state_bytes = (uint64_t)e->n_solid * Q * e->c_bytes;
    h.hist_bytes = (uint64_t)e->step * h.n_ucells * 8;
    const size_t total = sizeof h + h.pop_bytes + h.state_bytes + h.hist_bytes;
    if (needed) *needed = total;
    if (!buf) return NLOS_OK;
    if (cap < total) return fail(NLOS_ERR_ARG, "checkpoint buffer too small (%zu < %zu)", cap, total);
    HIP_TRY(hipStreamSynchronize(e->stream));
    char *p = static_cast<char *>(buf);
    std::memcpy(p, &h, sizeof h);
    p += sizeof h;
    HIP_TRY(hipMemcpy(p, e->curp(), h.pop_bytes, hipMemcpyDeviceToHost));
    p += h.pop_bytes;
    if (h.state_bytes) HIP_TRY(hipMemcpy(p, e->s_a.p, h.state_bytes, hipMemcpyDeviceToHost));
    p += h.state_bytes;
    if (h.hist_bytes) HIP_TRY(hipMemcpy(p, e->hist.p, h.hist_bytes, hipMemcpyDeviceToHost));
    return NLOS_OK;
}

NLOS_API int nlos_restore(nlos_engine *e, const void *buf, size_t len) {
    if (int rc = ready(e)) return rc;
    if (e->batch) return fail(NLOS_ERR_STATE, "scene belongs to a batch; it cannot be restored alone");
    if (!buf || len < sizeof(CkptHeader)) return fail(NLOS_ERR_ARG, "checkpoint too short");
    CkptHeader h;
    std::memcpy(&h, buf, sizeof h);
    if (std::memcmp(h.magic, "NLOSCKP1", 8) != 0) return fail(NLOS_ERR_ARG, "not a checkpoint");
    if (h.version != CKPT_VERSION)
        return fail(NLOS_ERR_ARG, "checkpoint format version %d is not %d", h.version, CKPT_VERSION);
    if (h.family != e->family) return fail(NLOS_ERR_ARG, "checkpoint was taken with the other solver family");
    if ((h.lay == LAY_AB) != (e->pattern == NLOS_PATTERN_TWO_ARRAY))
        return fail(NLOS_ERR_ARG, "checkpoint was taken with the other stream pattern");
    if (h.mode != e->mode || h.nx != e->g.nx || h.ny != e->g.ny || h.nz != e->g.nz || h.ncell != e->g.ncell ||
        h.n_solid != e->n_solid || h.n_ucells != (int64_t)e->probe_ucells.size())
        return fail(NLOS_ERR_ARG, "checkpoint does not match this engine's grid, geometry or probes");
    if (len < sizeof h + h.pop_bytes + h.state_bytes + h.hist_bytes) return fail(NLOS_ERR_ARG, "checkpoint truncated");
    const char *p = static_cast<const char *>(buf) + sizeof h;
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(hipMemcpy(e->curp(), p, h.pop_bytes, hipMemcpyHostToDevice));
    p += h.pop_bytes;
    if (h.state_bytes) HIP_TRY(hipMemcpy(e->s_a.p, p, h.state_bytes, hipMemcpyHostToDevice));
    p += h.state_bytes;
    e->step = 0;
    if (int rc = ensure_hist(e, h.step)) return rc;
    if (h.hist_bytes) HIP_TRY(hipMemcpy(e->hist.p, p, h.hist_bytes, hipMemcpyHostToDevice));
    e->step = h.step;
    e->lay = h.lay;
    e->field_valid = false;
    e->fb_valid = false;
    e->walls_valid = false;
    HIP_TRY(hipMemsetAsync(e->flag.p, 0, sizeof(int), e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return NLOS_OK;
}

// ---- measurement ------------------------------------------------------------------------------
NLOS_API int nlos_timing_enable(nlos_engine *e, int32_t on) {
    if (int rc = ready(e)) return rc;
    e->timing = on != 0;
    return NLOS_OK;
}

NLOS_API int nlos_timing_reset(nlos_engine *e) {
    if (int rc = ready(e)) return rc;
    if (int rc = collect_timing(e)) return rc;
    e->timed_launches = 0;
    e->timed_ms = 0.0;
    return NLOS_OK;
}

NLOS_API int nlos_timing_get(nlos_engine *e, nlos_timing *t) {
    if (int rc = ready(e)) return rc;
    if (!t) return fail(NLOS_ERR_ARG, "null timing");
    if (int rc = collect_timing(e)) return rc;
    t->launches = e->timed_launches;
    t->total_ms = e->timed_ms;
    t->avg_ms = e->timed_launches ? e->timed_ms / (double)e->timed_launches : 0.0;
    return NLOS_OK;
}

// Test hook (nlos_engine_test.h): FP64 full populations f* of every interior cell, AoS
// [cell][q], NaN on solid cells.
NLOS_API int nlos_debug_read_f(nlos_engine *e, double *out) {
    if (int rc = ready(e)) return rc;
    ScopedBuf tmp;
    HIP_TRY(tmp.reserve((size_t)e->n_interior() * Q * 8));
    HIP_TRY(by_mode(e->mode, [&](auto L) {
        return decltype(L)::gather_f(e->curp(), e->mask.as<uint8_t>(), e->g, e->lay, e->inv_scale, tmp.as<double>(), e->stream);
    }));
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(hipMemcpy(out, tmp.p, (size_t)e->n_interior() * Q * 8, hipMemcpyDeviceToHost));
    return NLOS_OK;
}

// Test hook (nlos_engine_test.h): overwrite one stored population of one interior cell (storage units are
// converted from the FP64 full population).  Used to force a non-finite value from the host.
NLOS_API int nlos_debug_poke(nlos_engine *e, int32_t i, int32_t j, int32_t k, int32_t q, double f) {
    if (int rc = ready(e)) return rc;
    if (i < 0 || i >= e->g.nx || j < 0 || j >= e->g.ny || k < 0 || k >= e->g.nz || q < 0 || q >= Q)
        return fail(NLOS_ERR_ARG, "poke outside grid");
    HIP_TRY(hipStreamSynchronize(e->stream));
    const long long idx = pop_index_host(e->mode, q, e->cell(i + 1, j + 1, k + 1), e->g.ncell);
    e->fb_valid = false;
    e->walls_valid = false;
    if (e->mode == M_F64) {
        HIP_TRY(hipMemcpy(static_cast<double *>(e->curp()) + idx, &f, 8, hipMemcpyHostToDevice));
    } else if (e->mode == M_F32) {
        const float v = (float)(f - W64[q]);
        HIP_TRY(hipMemcpy(static_cast<float *>(e->curp()) + idx, &v, 4, hipMemcpyHostToDevice));
    } else if (e->mode == M_B16) {
        // re-code the whole cell with the new value (see Mode in common.hpp)
        const long long N = e->g.ncell, c = e->cell(i + 1, j + 1, k + 1);
        char *base = static_cast<char *>(e->curp());
        float *rp = reinterpret_cast<float *>(base) + c;
        short *mp = reinterpret_cast<short *>(base) + 2 * N;
        signed char *ep = reinterpret_cast<signed char *>(base) + 40 * N + c;
        short m[Q] = {0};
        signed char E;
        float x0;
        for (int r = 1; r < Q; ++r) HIP_TRY(hipMemcpy(&m[r], mp + mant_index(r, c), 2, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&E, ep, 1, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&x0, rp, 4, hipMemcpyDeviceToHost));
        double x[Q], mx = 0.0;
        x[0] = x0;
        for (int r = 1; r < Q; ++r) x[r] = std::ldexp((double)m[r], E);
        x[q] = (f - W64[q]) / W64[q];
        for (int r = 1; r < Q; ++r) mx = std::max(mx, std::fabs(x[r]));
        int ex = 0;
        std::frexp((float)mx, &ex);
        ex = std::max(ex, B16_EMIN);
        for (int r = 1; r < Q; ++r) {
            const double v = std::nearbyint(std::ldexp(x[r], B16_BITS - ex));
            m[r] = (short)std::min(std::max(v, -32768.0), 32767.0);
        }
        E = (signed char)(ex - B16_BITS);
        x0 = (float)x[0];
        if (!std::isfinite(f) && q > 0) {
            // B16 mantissas have no NaN or Inf; the largest exponent decodes 32767 * 2^127 to
            // +Inf in FP32, so the next sweep forms a non-finite density as the other modes do
            m[q] = 32767;
            E = 127;
        }
        for (int r = 1; r < Q; ++r) HIP_TRY(hipMemcpy(mp + mant_index(r, c), &m[r], 2, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(ep, &E, 1, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(rp, &x0, 4, hipMemcpyHostToDevice));
    } else {
        const _Float16 v = (_Float16)((float)(f - W64[q]) * e->scale);
        HIP_TRY(hipMemcpy(static_cast<_Float16 *>(e->curp()) + idx, &v, 2, hipMemcpyHostToDevice));
    }
    return NLOS_OK;
}

namespace {
// Times `launch` repeated `reps` times on a private stream after one untimed launch; the
// stream and events are released on every return path.
template <class F>
int time_launches(int reps, double *avg_ms, F launch) {
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto done = [&](int rc) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (s) (void)hipStreamDestroy(s);
        return rc;
    };
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess)
        return done(fail(NLOS_ERR_HIP, "cannot create a stream or events for timing"));
    hipError_t r = launch(s);
    if (r == hipSuccess) r = hipEventRecord(e0, s);
    for (int k = 0; k < reps && r == hipSuccess; ++k) r = launch(s);
    if (r == hipSuccess) r = hipEventRecord(e1, s);
    if (r == hipSuccess) r = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (r == hipSuccess) r = hipEventElapsedTime(&ms, e0, e1);
    if (r != hipSuccess) return done(fail(NLOS_ERR_HIP, "timed launches failed: %s", hipGetErrorString(r)));
    *avg_ms = ms / reps;
    return done(NLOS_OK);
}
}  // namespace

NLOS_API int nlos_debug_pull_pattern(int32_t nx, int32_t ny, int32_t nz, int32_t reps, double *avg_ms) {
    if (nx < 1 || ny < 1 || nz < 1 || reps < 1 || !avg_ms) return fail(NLOS_ERR_ARG, "bad pattern request");
    const GridDesc g = make_grid(nx, ny, nz);
    const size_t bytes = (size_t)g.ncell * Q * 2;
    ScopedBuf a, b;
    if (a.reserve(bytes) != hipSuccess || b.reserve(bytes) != hipSuccess)
        return fail(NLOS_ERR_NOMEM, "cannot allocate two %.2f GB pattern arrays", bytes / 1e9);
    HIP_TRY(hipMemset(a.p, 0, bytes));
    return time_launches(reps, avg_ms, [&](hipStream_t s) { return launch_pull_pattern16(a.p, b.p, g, s); });
}

NLOS_API int nlos_debug_copy_kernel(size_t bytes, int32_t reps, double *avg_ms) {
    // one thread per 16 bytes: the grid stays below 2^31 workgroups up to 2^43 bytes
    if (bytes < 16 || bytes > ((size_t)1 << 40) || reps < 1 || !avg_ms) return fail(NLOS_ERR_ARG, "bad copy request");
    const long long n16 = (long long)(bytes / 16);
    ScopedBuf a, b;
    HIP_TRY(a.reserve((size_t)n16 * 16));
    HIP_TRY(b.reserve((size_t)n16 * 16));
    HIP_TRY(hipMemset(a.p, 0, (size_t)n16 * 16));
    return time_launches(reps, avg_ms, [&](hipStream_t s) { return launch_copy16(a.p, b.p, n16, s); });
}

// =========================================================================================
// Batched scenes.  Every scene is an ordinary engine (two-array pattern); the batch owns
// them and advances them together with multi-scene launches: per step one source launch
// (entries of all scenes, indexed by a device step counter), one ghost and one boundary
// solid launch over the concatenated halo / solid ranges, one sweep over all scenes' tiles
// (grid y = scene), and a counter increment.  Kernel arguments of every scene for
// the current state (phase 0) and the state one step later (phase 1) sit in device arrays,
// so a captured graph of two steps is replayed for every pair of steps.
struct nlos_batch {
    std::vector<nlos_engine *> scenes;
    int mode = M_F64;
    int device = 0;
    int n_cu = 0;
    hipStream_t stream = nullptr;
    dim3 block{256, 1, 1};                // sweep tile shape of every scene without its own width
    std::vector<int> tile_w;              // per-scene sweep tile widths (64, 128, 256), or empty
    DevBuf step_args[2], ghost_args[2], solid_args[2], source_args[2];
    DevBuf tile_table, halo_blocks, solid_blocks, ctr;
    DevBuf walk_items;                    // (scene, tile) work list of the persistent sweep
    long long total_items = 0;
    bool persistent = true;               // NLOS_BATCH_PERSISTENT=0: one workgroup per tile
    int walk_per_cu = 32;                 // persistent grid: workgroups per compute unit
    long long total_tiles = 0;            // sweep grid x: the largest scene's tile count
    long long total_halo = 0, total_solid = 0, max_entries = 0;   // table sizes
    DevBuf sched_off, sched_scene, sched_cells, sched_adds;
    long long h_ctr[2] = {0, 0};          // host staging of the counter and its run base
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<long long> baked;         // launch shape the graph was captured with
    long long step = 0;
    bool use_graph = true;                // NLOS_BATCH_GRAPH=0: launch every step directly
    bool timing = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double timed_ms = 0.0;
    long long timed_steps = 0;
};

namespace {

int batch_ready(const nlos_batch *b) {
    if (!b) return fail(NLOS_ERR_ARG, "null batch");
    return NLOS_OK;
}

int batch_settle(const nlos_batch *b) {
    HIP_TRY(hipStreamSynchronize(b->stream));
    return NLOS_OK;
}

// Sweep tile width of scene i: tiles of w x 256 / w cells, w = 64, 128 or 256.
int scene_tile_w(const nlos_batch *b, size_t i) { return b->tile_w.empty() ? (int)b->block.x : b->tile_w[i]; }

// Sweep arguments of scene i in the given phase for tile width w.
StepArgs multi_step_args(const nlos_batch *b, size_t i, int phase, int w) {
    StepArgs a = step_args(b->scenes[i], phase, dim3(w, 256 / w, 1));
    a.tile_w = w;
    return a;
}


long long stencil_bound(const nlos_engine *e) {
    if (!e->has_source) return 0;
    if (e->src.injection == NLOS_INJECT_TRILINEAR) return 8;
    const long long d = 2LL * e->src.stencil_radius + 1;
    return d * d * d;
}

void drop_graph(nlos_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Upload a host vector into a device buffer of at least its size; `moved` is set when the
// buffer had to be reallocated (its address is baked into a captured graph).
template <class T>
int upload_vec(nlos_batch *b, DevBuf &d, const std::vector<T> &v, bool &moved) {
    const size_t n = std::max<size_t>(v.size(), 1) * sizeof(T);
    if (n > d.bytes) {
        HIP_TRY(d.reserve(std::max(n, d.bytes * 2)));
        moved = true;
    }
    if (!v.empty()) HIP_TRY(hipMemcpyAsync(d.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, b->stream));
    return NLOS_OK;
}

// Per-scene argument arrays of both phases and the prefix arrays of tiles, halo cells and
// tracked solids, for the current scene states and launch shape.
int batch_prepare(nlos_batch *b, bool &moved) {
    const int n = (int)b->scenes.size();
    std::vector<StepArgs> st[2];
    std::vector<GhostArgs> gh[2];
    std::vector<SolidArgs> so[2];
    std::vector<SourceArgs> sr[2];
    std::vector<int> tiles;                // sweep tiles per scene
    std::vector<int2> halo, solid;         // work tables (scene, first item)
    std::vector<int2> walk;                // (scene, tile) of every sweep tile
    b->max_entries = 0;
    for (size_t si = 0; si < b->scenes.size(); ++si) {
        nlos_engine *e = b->scenes[si];
        HIP_TRY(hipStreamSynchronize(e->stream));
        if (int rc = sync_face_buffers(e)) return rc;
        HIP_TRY(hipStreamSynchronize(e->stream));
        for (int ph = 0; ph < 2; ++ph) {
            StepArgs a = multi_step_args(b, si, ph, scene_tile_w(b, si));
            a.step_ctr = b->ctr.as<long long>();
            a.hist = e->hist.as<double>();
            if (!a.hist) a.n_probe_cells = 0;
            st[ph].push_back(a);
            gh[ph].push_back(ghost_args(e, ph));
            so[ph].push_back(solid_args(e, ph));
            sr[ph].push_back(source_args(e, ph));
        }
        const int scene = (int)tiles.size();
        tiles.push_back(st[0].back().nblocks);
        if (b->persistent)
            for (int t = 0; t < tiles.back(); ++t) walk.push_back(make_int2(scene, t));
        for (long long h = 0; h < halo_cells(e); h += 256) halo.push_back(make_int2(scene, (int)h));
        for (int q = 0; q < e->n_solid; q += 256) solid.push_back(make_int2(scene, q));
        b->max_entries += stencil_bound(e);
    }
    for (int ph = 0; ph < 2; ++ph) {
        if (int rc = upload_vec(b, b->step_args[ph], st[ph], moved)) return rc;
        if (int rc = upload_vec(b, b->ghost_args[ph], gh[ph], moved)) return rc;
        if (int rc = upload_vec(b, b->solid_args[ph], so[ph], moved)) return rc;
        if (int rc = upload_vec(b, b->source_args[ph], sr[ph], moved)) return rc;
    }
    if (int rc = upload_vec(b, b->tile_table, tiles, moved)) return rc;
    if (int rc = upload_vec(b, b->walk_items, walk, moved)) return rc;
    b->total_items = (long long)walk.size();
    if (int rc = upload_vec(b, b->halo_blocks, halo, moved)) return rc;
    if (int rc = upload_vec(b, b->solid_blocks, solid, moved)) return rc;
    b->total_tiles = tiles.empty() ? 0 : *std::max_element(tiles.begin(), tiles.end());   // grid x
    b->total_halo = (long long)halo.size();
    b->total_solid = (long long)solid.size();
    (void)n;
    HIP_TRY(hipStreamSynchronize(b->stream));   // the host vectors above go out of scope
    return NLOS_OK;
}

SourceMulti batch_source_multi(const nlos_batch *b, int phase) {
    SourceMulti m{};
    m.scenes = b->source_args[phase].as<SourceArgs>();
    m.ctr = b->ctr.as<long long>();
    m.ctr_base = b->ctr.as<long long>() + 1;
    m.step_off = b->sched_off.as<long long>();
    m.entry_scene = b->sched_scene.as<int>();
    m.cells = b->sched_cells.as<long long>();
    m.adds = b->sched_adds.as<double>();
    return m;
}

// One step of every scene in the given phase.
int batch_launch_step(nlos_batch *b, int phase) {
    const int n = (int)b->scenes.size();
    const SourceMulti m = batch_source_multi(b, phase);
    HIP_TRY(by_mode(b->mode, [&](auto L) {
        using LL = decltype(L);
        hipError_t r = LL::source_multi(m, b->max_entries, b->stream);
        if (r == hipSuccess)
            r = LL::ghost_multi(b->ghost_args[phase].as<GhostArgs>(), b->halo_blocks.as<int2>(),
                                (int)b->total_halo, b->stream);
        if (r == hipSuccess)
            r = LL::solid_multi(b->solid_args[phase].as<SolidArgs>(), b->solid_blocks.as<int2>(),
                                (int)b->total_solid, b->stream);
        if (r == hipSuccess && b->persistent)
            r = LL::step_ab_walk(b->step_args[phase].as<StepArgs>(), b->walk_items.as<int2>(), (int)b->total_items,
                                 b->walk_per_cu * b->n_cu, b->stream);
        else if (r == hipSuccess)
            r = LL::step_ab_multi(b->step_args[phase].as<StepArgs>(), b->tile_table.as<int>(), n,
                                  (int)b->total_tiles, b->stream);
        return r;
    }));
    HIP_TRY(launch_count(b->ctr.as<long long>(), 1, b->stream));
    return NLOS_OK;
}

// The source schedule of steps step .. step+m-1 of every scene, CSR over steps.
int batch_schedule(nlos_batch *b, long long m, bool &moved) {
    // Each scene's entries for all m steps (host threads over scenes; every scene uses
    // only its own engine's buffers), then merged step-major.
    const size_t n = b->scenes.size();
    std::vector<std::vector<long long>> s_off(n);
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const size_t nt = std::min<size_t>(hw, n);
    auto work = [&](size_t t0) {
        for (size_t i = t0; i < n; i += nt) {
            nlos_engine *e = b->scenes[i];
            e->sched_cells.clear();
            e->sched_adds.clear();
            s_off[i].assign(1, 0);
            for (long long s = 0; s < m; ++s) {
                schedule_step(e, (double)(b->step + s) * e->dt);
                s_off[i].push_back((long long)e->sched_cells.size());
            }
        }
    };
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nt; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto &th : pool) th.join();
    std::vector<long long> off(1, 0), cells;
    std::vector<int> scene;
    std::vector<double> adds;
    for (long long s = 0; s < m; ++s) {
        for (size_t i = 0; i < n; ++i) {
            const nlos_engine *e = b->scenes[i];
            for (long long k = s_off[i][s]; k < s_off[i][s + 1]; ++k) {
                cells.push_back(e->sched_cells[k]);
                adds.push_back(e->sched_adds[k]);
                scene.push_back((int)i);
            }
        }
        off.push_back((long long)cells.size());
    }
    if (int rc = upload_vec(b, b->sched_off, off, moved)) return rc;
    if (int rc = upload_vec(b, b->sched_scene, scene, moved)) return rc;
    if (int rc = upload_vec(b, b->sched_cells, cells, moved)) return rc;
    if (int rc = upload_vec(b, b->sched_adds, adds, moved)) return rc;
    HIP_TRY(hipStreamSynchronize(b->stream));
    return NLOS_OK;
}

int batch_capture(nlos_batch *b) {
    drop_graph(b);
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = batch_launch_step(b, 0);
    if (!rc) rc = batch_launch_step(b, 1);
    hipGraph_t g = nullptr;
    const hipError_t he = hipStreamEndCapture(b->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (he != hipSuccess) return fail(NLOS_ERR_HIP, "graph capture: %s", hipGetErrorString(he));
    b->graph = g;
    HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return NLOS_OK;
}

}  // namespace

NLOS_API int nlos_batch_create(const nlos_config *cfgs, int32_t n, nlos_batch **out) {
    if (!cfgs || n <= 0 || !out) return fail(NLOS_ERR_ARG, "bad batch arguments");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        if (cfgs[i].precision != cfgs[0].precision) return fail(NLOS_ERR_ARG, "scenes of a batch share one precision");
        if (cfgs[i].family != NLOS_FAMILY_B) return fail(NLOS_ERR_UNSUPPORTED, "batches run NLOS_FAMILY_B scenes");
        if (cfgs[i].stream_pattern == NLOS_PATTERN_INPLACE)
            return fail(NLOS_ERR_UNSUPPORTED, "batched scenes use the two-array pattern");
    }
    nlos_batch *b = new nlos_batch();
    auto bail = [&](int rc) { nlos_batch_destroy(b); return rc; };
    b->mode = cfgs[0].precision;
    if (const char *x = std::getenv("NLOS_BATCH_BLOCK")) parse_block(x, b->block);
    if (const char *x = std::getenv("NLOS_BATCH_GRAPH")) b->use_graph = std::atoi(x) != 0;
    if (const char *x = std::getenv("NLOS_BATCH_PERSISTENT")) b->persistent = std::atoi(x) != 0;
    if (const char *x = std::getenv("NLOS_BATCH_WALK_PER_CU")) b->walk_per_cu = std::max(1, std::atoi(x));
    if (hipGetDevice(&b->device) != hipSuccess ||
        hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, b->device) != hipSuccess)
        return bail(fail(NLOS_ERR_HIP, "device query failed"));
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(NLOS_ERR_HIP, "stream creation failed"));
    if (b->ctr.reserve(2 * sizeof(long long)) != hipSuccess) return bail(fail(NLOS_ERR_NOMEM, "counter"));
    for (int i = 0; i < n; ++i) {
        nlos_config c = cfgs[i];
        c.stream_pattern = NLOS_PATTERN_TWO_ARRAY;
        nlos_engine *e = nullptr;
        if (int rc = nlos_create(&c, &e)) {
            const std::string msg = g_err;
            return bail(fail(rc, "scene %d: %s", i, msg.c_str()));
        }
        e->batch = b;
        b->scenes.push_back(e);
        // a scene is stepped by the batch only: its side stream would stay unused
        (void)hipEventDestroy(e->ev_fork);
        (void)hipEventDestroy(e->ev_join);
        (void)hipStreamDestroy(e->side_stream);
        e->ev_fork = e->ev_join = nullptr;
        e->side_stream = nullptr;
    }
    *out = b;
    return NLOS_OK;
}

NLOS_API int nlos_batch_destroy(nlos_batch *b) {
    if (!b) return NLOS_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (nlos_engine *e : b->scenes) {
        e->batch = nullptr;
        nlos_destroy(e);
    }
    DevBuf *bufs[] = {&b->step_args[0], &b->step_args[1], &b->ghost_args[0], &b->ghost_args[1],
                      &b->solid_args[0], &b->solid_args[1], &b->source_args[0], &b->source_args[1],
                      &b->tile_table, &b->walk_items, &b->halo_blocks, &b->solid_blocks, &b->ctr,
                      &b->sched_off, &b->sched_scene, &b->sched_cells, &b->sched_adds};
    for (DevBuf *d : bufs) d->release();
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return NLOS_OK;
}

NLOS_API int nlos_batch_size(const nlos_batch *b, int32_t *n) {
    if (int rc = batch_ready(b)) return rc;
    if (!n) return fail(NLOS_ERR_ARG, "null output");
    *n = (int32_t)b->scenes.size();
    return NLOS_OK;
}

NLOS_API int nlos_batch_scene(nlos_batch *b, int32_t i, nlos_engine **scene) {
    if (int rc = batch_ready(b)) return rc;
    if (!scene || i < 0 || i >= (int32_t)b->scenes.size()) return fail(NLOS_ERR_ARG, "scene index out of range");
    *scene = b->scenes[i];
    return NLOS_OK;
}

NLOS_API int nlos_batch_autotune(nlos_batch *b, int32_t reps, nlos_tune *result) {
    if (int rc = batch_ready(b)) return rc;
    if (reps < 1) return fail(NLOS_ERR_ARG, "reps must be positive");
    const int widths[] = {64, 128, 256};
    const size_t n = b->scenes.size();
    if (!b->ev0) HIP_TRY(hipEventCreate(&b->ev0));
    if (!b->ev1) HIP_TRY(hipEventCreate(&b->ev1));
    bool moved = false;
    if (int rc = batch_prepare(b, moved)) return rc;   // face records in sync
    int tried = 0;
    // Sweep time of the scenes `sel` with tile widths `w` (the fastest of reps launches):
    // the sweep alone, into the
    // array the next step overwrites (no probe rows, no face records), after one untimed
    // launch.  The argument and tile arrays are rewritten by the next batch_prepare.
    auto time_sweep = [&](const std::vector<int> &sel, const std::vector<int> &w, double &ms_out) -> int {
        std::vector<StepArgs> st;
        std::vector<int> tiles;
        std::vector<int2> walk;
        for (size_t k = 0; k < sel.size(); ++k) {
            StepArgs a = multi_step_args(b, (size_t)sel[k], 0, w[k]);
            a.n_probe_cells = 0;
            a.fb[0] = a.fb[1] = nullptr;
            st.push_back(a);
            tiles.push_back(a.nblocks);
            if (b->persistent)
                for (int t = 0; t < a.nblocks; ++t) walk.push_back(make_int2((int)k, t));
        }
        bool mv = false;
        if (int rc = upload_vec(b, b->step_args[0], st, mv)) return rc;
        if (int rc = upload_vec(b, b->tile_table, tiles, mv)) return rc;
        if (int rc = upload_vec(b, b->walk_items, walk, mv)) return rc;
        HIP_TRY(hipStreamSynchronize(b->stream));
        const int max_tiles = tiles.empty() ? 0 : *std::max_element(tiles.begin(), tiles.end());
        auto sweep = [&]() {   // in the form the steps use
            return by_mode(b->mode, [&](auto L) {
                if (b->persistent)
                    return decltype(L)::step_ab_walk(b->step_args[0].as<StepArgs>(), b->walk_items.as<int2>(),
                                                     (int)walk.size(), b->walk_per_cu * b->n_cu, b->stream);
                return decltype(L)::step_ab_multi(b->step_args[0].as<StepArgs>(), b->tile_table.as<int>(),
                                                  (int)sel.size(), max_tiles, b->stream);
            });
        };
        HIP_TRY(sweep());
        ms_out = 1e300;
        for (int r = 0; r < reps; ++r) {   // the fastest launch: the least disturbed one
            HIP_TRY(hipEventRecord(b->ev0, b->stream));
            HIP_TRY(sweep());
            HIP_TRY(hipEventRecord(b->ev1, b->stream));
            HIP_TRY(hipEventSynchronize(b->ev1));
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
            ms_out = std::min(ms_out, (double)ms);
        }
        ++tried;
        return NLOS_OK;
    };
    std::vector<int> all(n);
    for (size_t i = 0; i < n; ++i) all[i] = (int)i;
    // one width for every scene
    double best = 1e300;
    int best_w = (int)b->block.x;
    for (int w : widths) {
        double ms = 0.0;
        if (int rc = time_sweep(all, std::vector<int>(n, w), ms)) return rc;
        if (ms < best) { best = ms; best_w = w; }
    }
    // per class of row length (interior rows of 1, 2, 3 or more 64-cell waves): each class
    // timed alone with every width, then all scenes with the classes' choices
    std::vector<std::vector<int>> cls(4);
    for (size_t i = 0; i < n; ++i)
        cls[std::min(3, (b->scenes[i]->g.nx + 63) / 64 - 1)].push_back((int)i);
    int n_cls = 0;
    for (const auto &c : cls) n_cls += c.empty() ? 0 : 1;
    std::vector<int> per(n, best_w);
    bool use_per = false;
    double per_ms = 1e300;
    if (n_cls > 1) {
        for (const auto &c : cls) {
            if (c.empty()) continue;
            double cb = 1e300;
            int cw = best_w;
            for (int w : widths) {
                double ms = 0.0;
                if (int rc = time_sweep(c, std::vector<int>(c.size(), w), ms)) return rc;
                if (ms < cb) { cb = ms; cw = w; }
            }
            for (int i : c) per[(size_t)i] = cw;
        }
        // (a choice of one width for every class is a single width, already timed above)
        if (std::any_of(per.begin(), per.end(), [&](int w) { return w != per[0]; })) {
            if (int rc = time_sweep(all, per, per_ms)) return rc;
            use_per = per_ms < 0.99 * best;   // a gain above launch-to-launch noise
        }
    }
    if (use_per) {
        b->tile_w = per;
        best = per_ms;
    } else {
        b->tile_w.clear();
        b->block = dim3(best_w, 256 / best_w, 1);
    }
    drop_graph(b);
    b->baked.clear();
    if (result) {
        result->block_x = use_per ? 0 : (int32_t)best_w;
        result->block_y = use_per ? 0 : (int32_t)(256 / best_w);
        result->candidates = tried;
        result->per_scene = use_per ? 1 : 0;
        result->best_ms = best;
    }
    return NLOS_OK;
}

NLOS_API int nlos_batch_tile_widths(const nlos_batch *b, int32_t *widths, int32_t n) {
    if (int rc = batch_ready(b)) return rc;
    if (!widths || n != (int32_t)b->scenes.size()) return fail(NLOS_ERR_ARG, "need one width per scene");
    for (size_t i = 0; i < b->scenes.size(); ++i) widths[i] = scene_tile_w(b, i);
    return NLOS_OK;
}

NLOS_API int nlos_batch_set_tile_widths(nlos_batch *b, const int32_t *widths, int32_t n) {
    if (int rc = batch_ready(b)) return rc;
    if (!widths || n != (int32_t)b->scenes.size()) return fail(NLOS_ERR_ARG, "need one width per scene");
    for (int32_t i = 0; i < n; ++i)
        if (widths[i] != 64 && widths[i] != 128 && widths[i] != 256)
            return fail(NLOS_ERR_ARG, "tile width %d of scene %d is not 64, 128 or 256", widths[i], i);
    b->tile_w.assign(widths, widths + n);
    return NLOS_OK;
}

NLOS_API int nlos_batch_step(nlos_batch *b, int64_t n) {
    if (int rc = batch_ready(b)) return rc;
    if (n < 0) return fail(NLOS_ERR_ARG, "negative step count");
    if (n == 0) return NLOS_OK;
    for (nlos_engine *e : b->scenes) {
        if (e->step != b->step) return fail(NLOS_ERR_STATE, "scene stepped outside its batch");
        if (int rc = ensure_hist(e, e->step + n)) return rc;
    }
    if (b->timing) {
        if (!b->ev0) HIP_TRY(hipEventCreate(&b->ev0));
        if (!b->ev1) HIP_TRY(hipEventCreate(&b->ev1));
    }
    const long long chunk = 2048;
    for (long long base = 0; base < n; base += chunk) {
        const long long m = std::min<long long>(chunk, n - base);
        bool moved = false;
        if (int rc = batch_prepare(b, moved)) return rc;
        if (int rc = batch_schedule(b, m, moved)) return rc;
        b->h_ctr[0] = b->step;
        b->h_ctr[1] = b->step;
        HIP_TRY(hipMemcpyAsync(b->ctr.p, b->h_ctr, sizeof b->h_ctr, hipMemcpyHostToDevice, b->stream));
        HIP_TRY(hipStreamSynchronize(b->stream));
        // the tile shapes are read from the argument arrays; only the grids are baked
        const std::vector<long long> shape = {b->total_tiles, b->total_items, b->total_halo, b->total_solid,
                                              b->max_entries};
        if (moved || shape != b->baked) drop_graph(b);
        if (m >= 2 && !b->exec && b->use_graph) {
            if (int rc = batch_capture(b)) return rc;
            b->baked = shape;
        }
        if (b->timing) HIP_TRY(hipEventRecord(b->ev0, b->stream));
        for (long long r = 0; r < m / 2; ++r) {
            if (b->use_graph) {
                HIP_TRY(hipGraphLaunch(b->exec, b->stream));
            } else {
                // the two captured steps, launched one by one
                if (int rc = batch_launch_step(b, 0)) return rc;
                if (int rc = batch_launch_step(b, 1)) return rc;
            }
        }
        if (m & 1) {
            if (int rc = batch_launch_step(b, 0)) return rc;
        }
        if (b->timing) {
            HIP_TRY(hipEventRecord(b->ev1, b->stream));
            HIP_TRY(hipEventSynchronize(b->ev1));
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
            b->timed_ms += ms;
            b->timed_steps += m;
        }
        // phase bookkeeping: after m steps every scene is m steps on
        for (nlos_engine *e : b->scenes) {
            if (m & 1) {
                e->cur ^= 1;
                if (e->n_solid > 0) std::swap(e->s_a, e->s_b);
            }
            e->step += m;
            e->fb_valid = true;
        }
        b->step += m;
    }
    return NLOS_OK;
}

NLOS_API int nlos_batch_synchronize(nlos_batch *b) {
    if (int rc = batch_ready(b)) return rc;
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (nlos_engine *e : b->scenes)
        if (int rc = check_flag(e)) return rc;
    return NLOS_OK;
}

NLOS_API int nlos_batch_timing_enable(nlos_batch *b, int32_t on) {
    if (int rc = batch_ready(b)) return rc;
    b->timing = on != 0;
    if (b->timing) {
        b->timed_ms = 0.0;
        b->timed_steps = 0;
    }
    return NLOS_OK;
}

NLOS_API int nlos_batch_timing_get(nlos_batch *b, nlos_timing *t) {
    if (int rc = batch_ready(b)) return rc;
    if (!t) return fail(NLOS_ERR_ARG, "null timing");
    t->launches = b->timed_steps;
    t->total_ms = b->timed_ms;
    t->avg_ms = b->timed_steps ? b->timed_ms / (double)b->timed_steps : 0.0;
    return NLOS_OK;
}

// =========================================================================================
// Gradients by the discrete adjoint (k_adjoint): record, then backward over segments.
namespace {

int adjoint_supported(const nlos_engine *e) {
    if (e->pattern != NLOS_PATTERN_TWO_ARRAY)
        return fail(NLOS_ERR_UNSUPPORTED, "the adjoint needs the two-array stream pattern");
    if (e->cfg.nscbc_faces != 0 && e->n_solid > 0)
        return fail(NLOS_ERR_UNSUPPORTED, "the adjoint through NSCBC faces needs no solid cells on their boundary "
                    "planes (%d here)", e->n_solid);
    if (e->mode == M_F16 || e->mode == M_B16)
        return fail(NLOS_ERR_UNSUPPORTED, "the adjoint supports NLOS_F64 and NLOS_F32");
    if (e->family == NLOS_FAMILY_A) return fail(NLOS_ERR_UNSUPPORTED, "the adjoint is a NLOS_FAMILY_B feature");
    if (e->batch) return fail(NLOS_ERR_STATE, "scene belongs to a batch");
    return NLOS_OK;
}

}  // namespace

NLOS_API int nlos_record_forward(nlos_engine *e, int64_t n, int32_t every) {
    if (int rc = ready(e)) return rc;
    if (int rc = adjoint_supported(e)) return rc;
    if (n <= 0 || every <= 0) return fail(NLOS_ERR_ARG, "steps and checkpoint interval must be positive");
    for (DevBuf &b : e->ck) b.release();
    e->ck.clear();
    e->rec_start = e->step;
    e->rec_steps = 0;
    e->rec_every = every;
    for (long long done = 0; done < n; done += every) {
        DevBuf b;
        if (b.reserve(pop_bytes(e)) != hipSuccess) {
            e->rec_start = -1;
            return fail(NLOS_ERR_NOMEM, "cannot allocate checkpoint %zu", e->ck.size());
        }
        e->ck.push_back(b);   // owned by the engine from here on, whatever fails next
        HIP_TRY(hipMemcpyAsync(b.p, e->curp(), pop_bytes(e), hipMemcpyDeviceToDevice, e->stream));
        if (int rc = nlos_step(e, std::min<long long>(every, n - done), 0)) return rc;
        e->rec_steps += std::min<long long>(every, n - done);
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    return NLOS_OK;
}

NLOS_API int nlos_backward(nlos_engine *e, const double *dl_dprobe, int64_t max_steps, nlos_grad *grad) {
    if (int rc = ready(e)) return rc;
    if (int rc = adjoint_supported(e)) return rc;
    if (!grad) return fail(NLOS_ERR_ARG, "null gradient");
    if (e->rec_start < 0 || e->step != e->rec_start + e->rec_steps)
        return fail(NLOS_ERR_STATE, "no recorded forward ending at the current step (nlos_record_forward)");
    const long long s_end = e->rec_start + e->rec_steps;
    const size_t np = e->probes.size(), nu = e->probe_ucells.size();
    if (np > 0 && (!dl_dprobe || max_steps < s_end)) return fail(NLOS_ERR_ARG, "dl_dprobe must cover every recorded step");
    if (int rc = check_flag(e)) return rc;
    const int every = e->rec_every;
    const size_t pb = pop_bytes(e);
    const long long lam_n = e->g.ncell * (long long)Q;
    // dL/d(rho-1) per unique probe cell and recorded step
    std::vector<double> dl((size_t)e->rec_steps * std::max<size_t>(nu, 1), 0.0);
    for (size_t p = 0; p < np; ++p)
        for (long long s = 0; s < e->rec_steps; ++s)
            dl[(size_t)s * nu + e->probes[p].unique] += dl_dprobe[p * max_steps + e->rec_start + s];
    std::vector<DevBuf> store((size_t)every + 1);
    DevBuf lam[2], dl_dev, src_cells, src_adds, lam_src, vel;
    auto release = [&]() {
        for (DevBuf &b : store) b.release();
        lam[0].release(); lam[1].release(); dl_dev.release(); src_cells.release(); src_adds.release();
        lam_src.release();
        vel.release();
    };
    auto bail = [&](int rc) { release(); return rc; };
    for (DevBuf &b : store)
        if (b.reserve(pb) != hipSuccess) return bail(fail(NLOS_ERR_NOMEM, "cannot allocate %d recompute states", every + 1));
    // u of every step of a segment, recorded by the replay (3 values per cell and step)
    const size_t vel_step = (size_t)e->g.ncell * 3 * e->c_bytes;
    if (vel.reserve(vel_step * every) != hipSuccess) return bail(fail(NLOS_ERR_NOMEM, "cannot allocate velocities"));
    if (lam[0].reserve(lam_n * e->c_bytes) != hipSuccess || lam[1].reserve(lam_n * e->c_bytes) != hipSuccess)
        return bail(fail(NLOS_ERR_NOMEM, "cannot allocate adjoint arrays"));
    if (nu > 0) {
        if (dl_dev.reserve(dl.size() * 8) != hipSuccess) return bail(fail(NLOS_ERR_NOMEM, "probe adjoint"));
        if (hipMemcpy(dl_dev.p, dl.data(), dl.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(NLOS_ERR_HIP, "probe adjoint upload"));
    }
    auto hip = [&](hipError_t r, const char *what) {
        return r == hipSuccess ? NLOS_OK : fail(NLOS_ERR_HIP, "%s: %s", what, hipGetErrorString(r));
    };
    int rc = hip(by_mode(e->mode, [&](auto L) { return decltype(L)::fill_compute(lam[0].p, lam_n, e->stream); }),
                 "adjoint init");
    if (rc) return bail(rc);
    double g_pos[3] = {0.0, 0.0, 0.0}, g_amp = 0.0;
    std::vector<double> g_samp(e->samples.size(), 0.0);
    int cur = 0;
    const int nseg = (int)e->ck.size();
    for (int seg = nseg - 1; seg >= 0; --seg) {
        const long long s0 = e->rec_start + (long long)seg * every;
        const long long s1 = std::min<long long>(s0 + every, s_end);
        const int len = (int)(s1 - s0);
        // source entries of the segment with their derivatives
        std::vector<long long> off(1, 0);
        std::vector<double> unit;
        SourceDeriv d;
        e->sched_cells.clear();
        e->sched_adds.clear();
        for (long long n = s0; n < s1; ++n) {
            schedule_step(e, (double)n * e->dt, &d);
            off.push_back((long long)e->sched_cells.size());
            unit.push_back(waveform_unit(e, (double)n * e->dt));
        }
        const size_t ne = e->sched_cells.size();
        if (ne > 0) {
            if (src_cells.reserve(ne * 8) != hipSuccess || src_adds.reserve(ne * 8) != hipSuccess ||
                lam_src.reserve(ne * 8) != hipSuccess)
                return bail(fail(NLOS_ERR_NOMEM, "source entries"));
            if ((rc = hip(hipMemcpyAsync(src_cells.p, e->sched_cells.data(), ne * 8, hipMemcpyHostToDevice, e->stream), "upload")) ||
                (rc = hip(hipMemcpyAsync(src_adds.p, e->sched_adds.data(), ne * 8, hipMemcpyHostToDevice, e->stream), "upload")))
                return bail(rc);
        }
        // recompute the segment: store[k] = S_{s0+k} (after the source add of that step)
        if ((rc = hip(hipMemcpyAsync(store[0].p, e->ck[seg].p, pb, hipMemcpyDeviceToDevice, e->stream), "checkpoint copy")))
            return bail(rc);
        for (int k = 0; k < len; ++k) {
            SourceArgs sa = source_args(e, 0);
            sa.pop = store[k].p;
            sa.fb[0] = sa.fb[1] = nullptr;
            sa.cells = src_cells.as<long long>() + off[k];
            sa.adds = src_adds.as<double>() + off[k];
            sa.n = (int)(off[k + 1] - off[k]);
            // the adjoint reads these states whole, x-halo cells included: the replay keeps
            // the halo values in the lattice (no x-halo records)
            GhostArgs ga = ghost_args(e, 0);
            ga.pop = store[k].p;
            ga.fb[0] = ga.fb[1] = nullptr;
            ga.xh = nullptr;
            StepArgs st = step_args(e, 0, e->block);
            st.pop = store[k].p;
            st.pop_out = store[k + 1].p;
            st.n_probe_cells = 0;
            st.fb[0] = st.fb[1] = nullptr;
            st.xh = nullptr;
            st.vel = static_cast<char *>(vel.p) + vel_step * k;
            rc = hip(by_mode(e->mode, [&](auto L) {
                using LL = decltype(L);
                hipError_t r = LL::source(sa, LAY_AB, e->stream);
                if (r == hipSuccess) r = LL::ghost(ga, LAY_AB, e->stream);
                if (r == hipSuccess) r = LL::step_ab_vel(st, e->block, e->stream);
                return r;
            }), "recompute");
            if (rc) return bail(rc);
        }
        // adjoint steps back to the segment start
        for (int k = len - 1; k >= 0; --k) {
            const long long n = s0 + k;
            AdjArgs a{};
            a.state = store[k].p;
            a.lam_in = lam[cur].p;
            a.lam_out = lam[cur ^ 1].p;
            a.mask = e->mask.as<uint8_t>();
            a.wall_bits = e->wall_bits.as<uint32_t>();
            a.sig_x = e->sig[0].p;
            a.sig_y = e->sig[1].p;
            a.sig_z = e->sig[2].p;
            a.g = e->g;
            a.omega = e->omega;
            a.alpha = e->cfg.wall_reflect;
            a.scale = e->scale;
            a.inv_scale = e->inv_scale;
            a.probe_cells = e->probe_cells.as<long long>();
            a.n_probe_cells = (int)nu;
            a.dl_dp = nu ? dl_dev.as<double>() + (size_t)(n - e->rec_start) * nu : nullptr;
            a.nscbc_faces = e->cfg.nscbc_faces;
            a.vel = static_cast<const char *>(vel.p) + vel_step * k;
            const int cnt = (int)(off[k + 1] - off[k]);
            rc = hip(by_mode(e->mode, [&](auto L) {
                using LL = decltype(L);
                hipError_t r = LL::adjoint(a, e->stream);
                // the ghosts of this step, transposed onto their boundary cells (05:684-835)
                if (r == hipSuccess && e->cfg.nscbc_faces) r = LL::adjoint_ghost(a, e->stream);
                if (r == hipSuccess && cnt > 0)
                    r = LL::source_dot(lam[cur ^ 1].p, e->g.ncell, src_cells.as<long long>() + off[k], cnt,
                                       lam_src.as<double>() + off[k], e->stream);
                return r;
            }), "adjoint step");
            if (rc) return bail(rc);
            cur ^= 1;
        }
        // gradient of the segment's source terms (FP64, host)
        if (ne > 0) {
            std::vector<double> L(ne);
            if ((rc = hip(hipMemcpyAsync(L.data(), lam_src.p, ne * 8, hipMemcpyDeviceToHost, e->stream), "download")) ||
                (rc = hip(hipStreamSynchronize(e->stream), "sync")))
                return bail(rc);
            for (int k = 0; k < len; ++k) {
                const long long n = s0 + k;
                double dv = 0.0;
                for (long long q = off[k]; q < off[k + 1]; ++q) {
                    dv += L[q] * d.dval[q];
                    for (int ax = 0; ax < 3; ++ax) g_pos[ax] += L[q] * d.dc[3 * q + ax];
                }
                g_amp += dv * unit[k];
                if (e->src.waveform == NLOS_WAVE_SAMPLED) {
                    const long long si = std::llround((double)n * e->dt / e->dt);
                    if (si >= 0 && si < (long long)g_samp.size()) g_samp[(size_t)si] += dv * e->src.amp;
                }
            }
        }
    }
    if ((rc = hip(hipStreamSynchronize(e->stream), "sync"))) return bail(rc);
    release();
    for (int a = 0; a < 3; ++a) grad->d_position[a] = g_pos[a];
    grad->d_amp = g_amp;
    if (grad->d_samples)
        for (int64_t k = 0; k < grad->n_samples; ++k)
            grad->d_samples[k] = k < (int64_t)g_samp.size() ? g_samp[(size_t)k] : 0.0;
    return check_flag(e);
}
