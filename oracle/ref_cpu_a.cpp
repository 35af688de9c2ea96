// Test infrastructure, not part of the engine: a CPU restatement of the reference's
// family-A solver (01_acoustic_scattering.cpp, the step shared with 04_hpc_scalability.cpp)
// on one rank, in FP64 with contraction off.  Only tests/ use it, as the checker of the
// engine's legacy family-A mode.  Pinned against the probe traces of the compiled,
// unmodified 01 (tests/golden/ref01).
//
// Followed operation for operation:
//   velocity ordering A, inv_q, weights        01:17-27
//   get_feq                                     01:86-91
//   grid size, z forced to >= 3 cells           01:97-100
//   rest fill of f                              01:136-142
//   initial density f = feq(rho, 0) on fluid    01:166-205 (the pulse itself is the caller's)
//   sponge sigma (x/y faces, quadratic)         01:207-225
//   halo: x/y rest (no neighbour), z periodic   01:230-301 with periods {0,0,1} (01:110-112)
//   step_pull: collide at the source, bounce the
//     destination's own pre-collision value,
//     sponge of the destination, rho = sum f    01:303-359
//   probe cell (int)((x-min)/dx - 0.5) + 1,
//     rho-1 recorded before each step           01:448-479, 01:507-510
//   04's line source, added before the halo
//     exchange of every step                    04:212-235, 04:369-388
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/nlos_engine.h"

namespace {

constexpr int Q = 19;
constexpr int EXA[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int EYA[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int EZA[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
constexpr double WA[Q] = {1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
                          1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                          1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
constexpr int INVA[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};

double feq(int k, double r, double ux, double uy, double uz) {
    const double cu = EXA[k] * ux + EYA[k] * uy + EZA[k] * uz;
    const double u2 = ux * ux + uy * uy + uz * uz;
    return WA[k] * r * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * u2);
}

}  // namespace

struct oracle_a {
    int nx, ny, nz, nxt, nyt, nzt;
    double xmin, xmax, ymin, ymax, zmin, dx, tau;
    std::vector<double> f, fn, rho, sigma;
    std::vector<uint8_t> solid;
    std::vector<size_t> probe_cells;   // index into the halo-padded grid, or SIZE_MAX
    std::vector<std::vector<double>> hist;
    long long steps = 0;
    nlos_line_source line{};

    size_t idx(int i, int j, int k) const { return (size_t)i + (size_t)j * nxt + (size_t)k * nxt * nyt; }

    void record() {
        for (size_t p = 0; p < probe_cells.size(); ++p)
            hist[p].push_back(probe_cells[p] == SIZE_MAX ? 0.0 : rho[probe_cells[p]] - 1.0);
    }

    // apply_source of 04 (04:212-235): t_phys = (s - step0) * dt as its loops count s
    void apply_line_source() {
        const nlos_line_source &L = line;
        if (L.amp == 0.0) return;
        const double t_phys = (double)(steps - L.step0) * L.dt;
        const double dt_phys = L.dt;
#pragma omp parallel for collapse(3)
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    double x = xmin + (double)(i - 1) * dx + 0.5 * dx;
                    double y = ymin + (double)(j - 1) * dx + 0.5 * dx;
                    double dist2 = (x - L.x0) * (x - L.x0) + (y - L.y0) * (y - L.y0);
                    if (dist2 > L.r2_max) continue;
                    double spatial = std::exp(-std::log(2.0) * dist2 / L.b);
                    double source_val = L.amp * spatial * std::sin(L.omega * t_phys);
                    size_t c = idx(i, j, k);
                    if (!solid[c])
                        for (int q = 0; q < Q; ++q) f[c * Q + q] += WA[q] * source_val * dt_phys;
                }
    }

    void halos() {
        double f0[Q];
        for (int q = 0; q < Q; ++q) f0[q] = feq(q, 1.0, 0, 0, 0);
        auto rest = [&](size_t c) { for (int q = 0; q < Q; ++q) f[c * Q + q] = f0[q]; };
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j) { rest(idx(0, j, k)); rest(idx(nx + 1, j, k)); }
        for (int k = 1; k <= nz; ++k)
            for (int i = 0; i <= nx + 1; ++i) { rest(idx(i, 0, k)); rest(idx(i, ny + 1, k)); }
        for (int j = 0; j <= ny + 1; ++j)
            for (int i = 0; i <= nx + 1; ++i)
                for (int q = 0; q < Q; ++q) {
                    f[idx(i, j, 0) * Q + q] = f[idx(i, j, nz) * Q + q];
                    f[idx(i, j, nz + 1) * Q + q] = f[idx(i, j, 1) * Q + q];
                }
    }

    void step() {
#pragma omp parallel for collapse(3)
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    const size_t dst = idx(i, j, k);
                    if (solid[dst]) continue;
                    const double s = sigma[dst];
                    const bool in_sponge = s > 1e-6;
                    for (int q = 0; q < Q; ++q) {
                        const size_t src = idx(i - EXA[q], j - EYA[q], k - EZA[q]);
                        if (solid[src]) {
                            fn[dst * Q + q] = f[dst * Q + INVA[q]];
                            continue;
                        }
                        double sr = 0.0, sux = 0.0, suy = 0.0, suz = 0.0;
                        for (int p = 0; p < Q; ++p) {
                            const double v = f[src * Q + p];
                            sr += v;
                            sux += v * EXA[p];
                            suy += v * EYA[p];
                            suz += v * EZA[p];
                        }
                        if (sr == 0.0) sr = 1.0;
                        sux /= sr;
                        suy /= sr;
                        suz /= sr;
                        const double fe = feq(q, sr, sux, suy, suz);
                        const double post = f[src * Q + q] - (f[src * Q + q] - fe) / tau;
                        fn[dst * Q + q] = in_sponge ? (1.0 - s) * post + s * feq(q, 1.0, 0.0, 0.0, 0.0) : post;
                    }
                    double sum = 0.0;
                    for (int p = 0; p < Q; ++p) sum += fn[dst * Q + p];
                    rho[dst] = sum;
                }
        f.swap(fn);
    }
};

extern "C" {

oracle_a *oracle_a_create(const nlos_config *cfg) {
    oracle_a *o = new oracle_a();
    o->xmin = cfg->x_min; o->xmax = cfg->x_max;
    o->ymin = cfg->y_min; o->ymax = cfg->y_max;
    o->zmin = cfg->z_min;
    o->dx = cfg->dx;
    o->tau = cfg->tau;
    o->nx = (int)((cfg->x_max - cfg->x_min) / cfg->dx + 0.5);
    o->ny = (int)((cfg->y_max - cfg->y_min) / cfg->dx + 0.5);
    o->nz = (int)((cfg->z_max - cfg->z_min) / cfg->dx + 0.5);
    if (o->nz < 3) o->nz = 3;
    o->nxt = o->nx + 2; o->nyt = o->ny + 2; o->nzt = o->nz + 2;
    const size_t n = (size_t)o->nxt * o->nyt * o->nzt;
    o->f.assign(n * Q, 0.0);
    o->fn.assign(n * Q, 0.0);
    o->rho.assign(n, 1.0);
    o->sigma.assign(n, 0.0);
    o->solid.assign(n, 0);
    for (size_t c = 0; c < n; ++c)
        for (int q = 0; q < Q; ++q) o->f[c * Q + q] = feq(q, 1.0, 0, 0, 0);
    // sponge of 01:207-225 on the x/y faces, thickness sponge_width, strength sponge_strength
    const double T = cfg->sponge_width, S = cfg->sponge_strength;
    for (int k = 0; k < o->nzt; ++k)
        for (int j = 0; j < o->nyt; ++j)
            for (int i = 0; i < o->nxt; ++i) {
                const double x = o->xmin + ((i - 1) + 0.5) * o->dx;
                const double y = o->ymin + ((j - 1) + 0.5) * o->dx;
                double dxs = 0.0, dys = 0.0;
                if (x < o->xmin + T) dxs = (o->xmin + T) - x;
                else if (x > o->xmax - T) dxs = x - (o->xmax - T);
                if (y < o->ymin + T) dys = (o->ymin + T) - y;
                else if (y > o->ymax - T) dys = y - (o->ymax - T);
                const double dm = std::max(dxs, dys);
                if (dm > 0.0) {
                    const double r = dm / T;
                    o->sigma[o->idx(i, j, k)] = S * r * r;
                }
            }
    return o;
}

void oracle_a_destroy(oracle_a *o) { delete o; }

void oracle_a_dims(const oracle_a *o, int32_t *d) { d[0] = o->nx; d[1] = o->ny; d[2] = o->nz; }

// Interior solid flags, x fastest.  01 evaluates its geometry on every cell, halos included
// (01:170-180); with z periodic the z-halo planes hold the geometry of planes nz and 1, and
// the x/y halo cells are fluid (the caller's geometry does not extend past the box).
void oracle_a_set_mask(oracle_a *o, const uint8_t *m) {
    size_t n = 0;
    for (int k = 1; k <= o->nz; ++k)
        for (int j = 1; j <= o->ny; ++j)
            for (int i = 1; i <= o->nx; ++i) {
                const size_t c = o->idx(i, j, k);
                o->solid[c] = m[n++] ? 1 : 0;
                if (o->solid[c]) o->sigma[c] = 0.0;   // 01 sets sigma on fluid cells only
            }
    for (int j = 1; j <= o->ny; ++j)
        for (int i = 1; i <= o->nx; ++i) {
            o->solid[o->idx(i, j, 0)] = o->solid[o->idx(i, j, o->nz)];
            o->solid[o->idx(i, j, o->nz + 1)] = o->solid[o->idx(i, j, 1)];
        }
}

// Interior sponge override, x fastest.
void oracle_a_set_sponge(oracle_a *o, const double *s) {
    size_t n = 0;
    for (int k = 1; k <= o->nz; ++k)
        for (int j = 1; j <= o->ny; ++j)
            for (int i = 1; i <= o->nx; ++i) o->sigma[o->idx(i, j, k)] = s[n++];
}

void oracle_a_get_sponge(const oracle_a *o, double *s) {
    size_t n = 0;
    for (int k = 1; k <= o->nz; ++k)
        for (int j = 1; j <= o->ny; ++j)
            for (int i = 1; i <= o->nx; ++i) s[n++] = o->sigma[o->idx(i, j, k)];
}

// rho and f = feq(rho, 0) on fluid interior cells (01:196-205).
void oracle_a_set_density(oracle_a *o, const double *rho) {
    size_t n = 0;
    for (int k = 1; k <= o->nz; ++k)
        for (int j = 1; j <= o->ny; ++j)
            for (int i = 1; i <= o->nx; ++i) {
                const size_t c = o->idx(i, j, k);
                const double r = rho[n++];
                if (o->solid[c]) continue;
                o->rho[c] = r;
                for (int q = 0; q < Q; ++q) o->f[c * Q + q] = o->fn[c * Q + q] = feq(q, r, 0.0, 0.0, 0.0);
            }
}

int oracle_a_set_probes(oracle_a *o, const double *xyz, int32_t n) {
    o->probe_cells.clear();
    o->hist.assign((size_t)n, {});
    for (int p = 0; p < n; ++p) {
        const int gx = (int)((xyz[3 * p] - o->xmin) / o->dx - 0.5);
        const int gy = (int)((xyz[3 * p + 1] - o->ymin) / o->dx - 0.5);
        const int gz = (int)((xyz[3 * p + 2] - o->zmin) / o->dx - 0.5);
        const int lx = gx + 1, ly = gy + 1, lz = gz + 1;
        const bool in = lx >= 1 && lx <= o->nx && ly >= 1 && ly <= o->ny && lz >= 1 && lz <= o->nz;
        o->probe_cells.push_back(in ? o->idx(lx, ly, lz) : SIZE_MAX);
    }
    return 0;
}

// n steps: record, halos, step (01:507-527).  The first call also records nothing extra;
// read_probes returns steps + 1 values per probe (s = 0 .. steps).
void oracle_a_run(oracle_a *o, long long n) {
    for (long long s = 0; s < n; ++s) {
        if (o->steps == 0 && o->hist.size() && o->hist[0].empty()) o->record();
        o->apply_line_source();
        o->halos();
        o->step();
        ++o->steps;
        o->record();
    }
}

void oracle_a_set_line_source(oracle_a *o, const nlos_line_source *s) { o->line = *s; }

long long oracle_a_read_probes(const oracle_a *o, double *out) {
    long long m = 0;
    for (size_t p = 0; p < o->hist.size(); ++p) {
        m = (long long)o->hist[p].size();
        std::memcpy(out + p * m, o->hist[p].data(), (size_t)m * 8);
    }
    return m;
}

// Interior f, [cell][q] in ordering A, x fastest.
void oracle_a_read_f(const oracle_a *o, double *out) {
    size_t n = 0;
    for (int k = 1; k <= o->nz; ++k)
        for (int j = 1; j <= o->ny; ++j)
            for (int i = 1; i <= o->nx; ++i, ++n)
                for (int q = 0; q < Q; ++q) out[n * Q + q] = o->f[o->idx(i, j, k) * Q + q];
}

// Interior rho (1 on solid cells), x fastest.
void oracle_a_read_rho(const oracle_a *o, double *out) {
    size_t n = 0;
    for (int k = 1; k <= o->nz; ++k)
        for (int j = 1; j <= o->ny; ++j)
            for (int i = 1; i <= o->nx; ++i) out[n++] = o->rho[o->idx(i, j, k)];
}

}  // extern "C"
