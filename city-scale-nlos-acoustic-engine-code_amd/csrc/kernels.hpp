// Device kernels of the D3Q19 engine, templated on the storage mode.  Included once by
// each mode's translation unit; the F64 unit is compiled with -ffp-contract=off so that
// its arithmetic rounds exactly like the reference's CPU code.
//
// Reference semantics of one time step s (05_sensor_illusion.cpp:1079-1093):
//   apply_source(t)     f*(x) += w_q*A*ww(x) on fluid stencil cells      05:576-667
//   exchange_halos()    halo ghost = NSCBC(f*(I)) on open faces, else w  05:671-904
//   step_collision_bgk  f_in_q(x) = solid(x-e_q) ? f*_{inv q}(x) : f*_q(x-e_q);
//                       BGK collide, sponge blend                         05:906-972
// f* is the post-collision state of the previous step.  The engine keeps f* in one
// array in the in-place A-A arrangement (see Layout in common.hpp) and alternates
// odd-type and even-type sweeps; ghosts, source and tracked solid cells are written into
// whichever layout is current.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "launch.hpp"

namespace nlos {

template <int MODE> struct Tr;
template <> struct Tr<M_F64> { using S = double; using C = double; };
template <> struct Tr<M_F32> { using S = float; using C = float; };
template <> struct Tr<M_F16> { using S = _Float16; using C = float; };

template <int MODE>
__device__ __forceinline__ typename Tr<MODE>::C ld(const void *base, long long idx, float inv_scale) {
    using S = typename Tr<MODE>::S;
    const S v = static_cast<const S *>(base)[idx];
    if constexpr (MODE == M_F16) return static_cast<float>(v) * inv_scale;
    else return v;
}

template <int MODE>
__device__ __forceinline__ void st(void *base, long long idx, typename Tr<MODE>::C v, float scale) {
    using S = typename Tr<MODE>::S;
    if constexpr (MODE == M_F16) static_cast<S *>(base)[idx] = static_cast<S>(v * scale);
    else static_cast<S *>(base)[idx] = v;
}

// Rest value of a stored population: w_q for full populations, 0 for deviations.
template <int MODE>
__device__ __forceinline__ typename Tr<MODE>::C rest(int q) {
    if constexpr (MODE == M_F64) return W64[q];
    else return 0.0f;
}

template <class C>
__device__ __forceinline__ C wq(int q) { return static_cast<C>(W64[q]); }

// Neighbour offset of direction q.
__device__ __forceinline__ long long nb_off(const GridDesc &g, int q) {
    return (long long)EX[q] + (long long)EY[q] * g.px + (long long)EZ[q] * g.plane;
}

// e_q . u written per direction so that no product with a zero component is formed; it
// equals the reference's ex*ux + ey*uy + ez*uz (05:288) up to the sign of a zero result,
// which the equilibrium does not see.
template <class C>
__device__ __forceinline__ C cu_of(int q, C ux, C uy, C uz) {
    switch (q) {
    case 0: return C(0);
    case 1: return ux;
    case 2: return -ux;
    case 3: return uy;
    case 4: return -uy;
    case 5: return uz;
    case 6: return -uz;
    case 7: return ux + uy;
    case 8: return ux - uy;
    case 9: return uy - ux;
    case 10: return -(ux + uy);
    case 11: return ux + uz;
    case 12: return ux - uz;
    case 13: return uz - ux;
    case 14: return -(ux + uz);
    case 15: return uy + uz;
    case 16: return uy - uz;
    case 17: return uz - uy;
    default: return -(uy + uz);
    }
}

// ---- full-population FP64 arithmetic in the reference's order -----------------------

// Zeroth and first moments summed in q order 0..18 (05:940-947, 05:549-556); terms with a
// zero velocity component are left out, which does not change any rounding.
__device__ __forceinline__ void moments64(const double *f, double &r, double &jx, double &jy, double &jz) {
    r = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8] + f[9] + f[10] + f[11] +
        f[12] + f[13] + f[14] + f[15] + f[16] + f[17] + f[18];
    jx = f[1] - f[2] + f[7] + f[8] - f[9] - f[10] + f[11] + f[12] - f[13] - f[14];
    jy = f[3] - f[4] + f[7] - f[8] + f[9] - f[10] + f[15] + f[16] - f[17] - f[18];
    jz = f[5] - f[6] + f[11] - f[12] + f[13] - f[14] + f[15] - f[16] + f[17] - f[18];
}

// LBMSolver::feq (05:286-291).
__device__ __forceinline__ double feq64(int q, double r, double ux, double uy, double uz, double u2) {
    const double cu = cu_of<double>(q, ux, uy, uz);
    return W64[q] * r * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * u2);
}

// Collision of step_collision_bgk (05:940-968).  Returns rho (05:953).
__device__ __forceinline__ double collide64(double *f, double omega, double sponge) {
    double r, jx, jy, jz;
    moments64(f, r, jx, jy, jz);
    if (r <= 0.0) r = 1e-12;
    const double ux = jx / r, uy = jy / r, uz = jz / r;
    const double u2 = ux * ux + uy * uy + uz * uz;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const double fe = feq64(q, r, ux, uy, uz, u2);
        double fp = f[q] - omega * (f[q] - fe);
        if (sponge > 1e-12) fp = (1.0 - sponge) * fp + sponge * W64[q];
        f[q] = fp;
    }
    return r;
}

// ---- deviation arithmetic (FP32) ------------------------------------------------------

__device__ __forceinline__ void moments32(const float *g, float &dr, float &jx, float &jy, float &jz) {
    dr = ((((g[0] + g[1]) + (g[2] + g[3])) + ((g[4] + g[5]) + (g[6] + g[7]))) +
          (((g[8] + g[9]) + (g[10] + g[11])) + ((g[12] + g[13]) + (g[14] + g[15])))) +
         ((g[16] + g[17]) + g[18]);
    jx = ((g[1] - g[2]) + (g[7] - g[10])) + ((g[8] - g[9]) + (g[11] - g[14])) + (g[12] - g[13]);
    jy = ((g[3] - g[4]) + (g[7] - g[10])) + ((g[9] - g[8]) + (g[15] - g[18])) + (g[16] - g[17]);
    jz = ((g[5] - g[6]) + (g[11] - g[14])) + ((g[13] - g[12]) + (g[15] - g[18])) + (g[17] - g[16]);
}

// g_eq = feq - w = w*(drho + rho*(3cu + 4.5cu^2 - 1.5u^2)).
__device__ __forceinline__ float geq32(int q, float dr, float r, float ux, float uy, float uz, float u2c) {
    const float cu = cu_of<float>(q, ux, uy, uz);
    return wq<float>(q) * (dr + r * (cu * (3.0f + 4.5f * cu) - u2c));
}

// BGK collision in deviation form with the sponge as a factor (1 - sigma) on g
// ((1-s)f + s*w - w = (1-s)g).  Returns drho.
__device__ __forceinline__ float collide32(float *g, float omega, float keep) {
    float dr, jx, jy, jz;
    moments32(g, dr, jx, jy, jz);
    float r = 1.0f + dr;
    if (r <= 0.0f) { r = 1e-12f; dr = r - 1.0f; }
    const float ir = 1.0f / r;
    const float ux = jx * ir, uy = jy * ir, uz = jz * ir;
    const float u2c = 1.5f * (ux * ux + uy * uy + uz * uz);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const float ge = geq32(q, dr, r, ux, uy, uz, u2c);
        g[q] = keep * (g[q] + omega * (ge - g[q]));
    }
    return dr;
}

// Reflection of a population bounced off a wall: the deviation from rest is scaled by
// alpha (alpha = 1 is the reference's plain bounce-back and leaves the value untouched).
template <int MODE>
__device__ __forceinline__ typename Tr<MODE>::C reflect(int q, typename Tr<MODE>::C v, double alpha) {
    if constexpr (MODE == M_F64) return W64[q] + alpha * (v - W64[q]);
    else return static_cast<float>(alpha) * v;
}

// ---- fused collide-stream sweep ---------------------------------------------------------

__device__ __forceinline__ int remap_block(int b, int nb) {
    // Workgroups are dealt to the 8 XCDs round-robin; give each XCD a contiguous run of
    // tiles so that the lines a tile shares with its neighbours sit in the same L2.
    const int qq = nb >> 3, rr = nb & 7, xcd = b & 7, idx = b >> 3;
    return xcd < rr ? xcd * (qq + 1) + idx : rr * (qq + 1) + (xcd - rr) * qq + idx;
}

__device__ __forceinline__ int probe_slot(const long long *cells, int n, long long c) {
    int lo = 0, hi = n - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const long long v = cells[mid];
        if (v == c) return mid;
        if (v < c) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

template <int MODE, bool ODD, bool CAPTURE>
__global__ __launch_bounds__(256) void k_step(StepArgs a) {
    using C = typename Tr<MODE>::C;
    int b = blockIdx.x;
    if (a.xcd_remap) b = remap_block(b, gridDim.x);
    const int bi = b % a.tiles_x;
    const int rest_b = b / a.tiles_x;
    const int bj = rest_b % a.tiles_y;
    const int k = rest_b / a.tiles_y + 1;
    const int i = bi * (int)blockDim.x + (int)threadIdx.x + 1;
    const int j = bj * (int)blockDim.y + (int)threadIdx.y + 1;
    if (i > a.g.nx || j > a.g.ny) return;

    const long long N = a.g.ncell;
    const long long c = (long long)i + (long long)a.g.px * ((long long)j + (long long)a.g.py * k);
    const uint8_t m = a.mask[c];
    if (m & MK_SOLID) return;

    // Solid-neighbour bits: bit q set when c + e_q is solid.
    uint32_t sb = 0;
    const bool need_nb = (m & MK_NEAR) && (ODD || a.alpha != 1.0);
    if (need_nb) {
#pragma unroll
        for (int q = 1; q < Q; ++q)
            if (a.mask[c + nb_off(a.g, q)] & MK_SOLID) sb |= 1u << q;
    }

    C f[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const bool from_wall = (sb >> INV[q]) & 1u;     // c - e_q is solid
        if constexpr (ODD) {
            // LAY_P: f*_q(c - e_q) sits in slot INV[q] of c - e_q; a bounced f*_{INV q}(c)
            // sits in slot q of c.
            const long long idx = from_wall ? (long long)q * N + c
                                            : (long long)INV[q] * N + c - nb_off(a.g, q);
            f[q] = ld<MODE>(a.pop, idx, a.inv_scale);
        } else {
            // LAY_Q: both the streamed and the bounced population sit in slot q of c.
            f[q] = ld<MODE>(a.pop, (long long)q * N + c, a.inv_scale);
        }
        if (from_wall && a.alpha != 1.0) f[q] = reflect<MODE>(q, f[q], a.alpha);
    }

    const C *sx = static_cast<const C *>(a.sig_x);
    const C *sy = static_cast<const C *>(a.sig_y);
    const C *sz = static_cast<const C *>(a.sig_z);
    // sigma(min d) = max over axes of sigma(d_axis): the ramp is monotone (05:491-506).
    const C sig = fmax(sx[i], fmax(sy[j], sz[k]));

    double dval;   // rho - 1
    bool finite;
    if constexpr (MODE == M_F64) {
        const double r = collide64(f, a.omega, sig);
        dval = r - 1.0;
        finite = isfinite(r);
    } else {
        const float dr = collide32(f, (float)a.omega, 1.0f - sig);
        dval = dr;
        finite = isfinite(dr);
    }
    if (!finite) atomicOr(a.flag, 1);
    if (m & MK_PROBE) {
        const int u = probe_slot(a.probe_cells, a.n_probe_cells, c);
        if (u >= 0) a.probe_row[u] = dval;
    }
    if constexpr (CAPTURE) {
        a.field[(long long)(i - 1) + (long long)a.g.nx * ((long long)(j - 1) + (long long)a.g.ny * (k - 1))] =
            (float)dval;
    }

#pragma unroll
    for (int q = 0; q < Q; ++q) {
        if constexpr (ODD) {
            // -> LAY_Q: slot q of c + e_q, or slot INV[q] of c when c + e_q is solid.
            const bool to_wall = (sb >> q) & 1u;
            const long long idx = to_wall ? (long long)INV[q] * N + c
                                          : (long long)q * N + c + nb_off(a.g, q);
            st<MODE>(a.pop, idx, f[q], a.scale);
        } else {
            // -> LAY_P: slot INV[q] of c.
            st<MODE>(a.pop, (long long)INV[q] * N + c, f[q], a.scale);
        }
    }
}

// ---- reading f* of an interior cell in the current layout ---------------------------------

template <int MODE, int LAY>
__device__ __forceinline__ typename Tr<MODE>::C read_post(const void *pop, const uint8_t *mask,
                                                          const GridDesc &g, long long c, int q,
                                                          float inv_scale) {
    const long long N = g.ncell;
    if constexpr (LAY == LAY_P) {
        return ld<MODE>(pop, (long long)INV[q] * N + c, inv_scale);
    } else {
        const long long t = c + nb_off(g, q);
        const long long idx = (mask[t] & MK_SOLID) ? (long long)INV[q] * N + c : (long long)q * N + t;
        return ld<MODE>(pop, idx, inv_scale);
    }
}

// ---- ghost layer (05:671-904 at one rank) --------------------------------------------------

template <int MODE, int LAY>
__global__ __launch_bounds__(256) void k_ghost(GhostArgs a) {
    using C = typename Tr<MODE>::C;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n_halo) return;
    const GridDesc &g = a.g;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    int i, j, k;
    // Halo enumeration: the two z layers, then the two y rows of every interior k, then
    // the two x columns of every interior (j,k).
    const long long zpl = (long long)(nx + 2) * (ny + 2);
    const long long yrow = (long long)(nx + 2) * nz;
    if (t < 2 * zpl) {
        k = t < zpl ? 0 : nz + 1;
        const long long r = t < zpl ? t : t - zpl;
        j = (int)(r / (nx + 2));
        i = (int)(r % (nx + 2));
    } else if ((t -= 2 * zpl) < 2 * yrow) {
        j = t < yrow ? 0 : ny + 1;
        const long long r = t < yrow ? t : t - yrow;
        k = (int)(r / (nx + 2)) + 1;
        i = (int)(r % (nx + 2));
    } else {
        t -= 2 * yrow;
        const long long xcol = (long long)ny * nz;
        i = t < xcol ? 0 : nx + 1;
        const long long r = t < xcol ? t : t - xcol;
        k = (int)(r / ny) + 1;
        j = (int)(r % ny) + 1;
    }
    const long long N = g.ncell;
    const long long h = (long long)i + (long long)g.px * ((long long)j + (long long)g.py * k);
    if (a.mask[h] & MK_SOLID) return;

    const int oi = i == 0 ? -1 : (i == nx + 1 ? 1 : 0);
    const int oj = j == 0 ? -1 : (j == ny + 1 ? 1 : 0);
    const int ok = k == 0 ? -1 : (k == nz + 1 ? 1 : 0);
    const int nout = (oi != 0) + (oj != 0) + (ok != 0);
    int axis = 0, face = 0;
    bool ns = false;
    if (nout == 1) {
        axis = oi ? 0 : (oj ? 1 : 2);
        face = 2 * axis + ((oi + oj + ok) > 0 ? 1 : 0);
        ns = (a.nscbc_faces >> face) & 1;
    }

    // Boundary state of an NSCBC face cell (05:689-699 and its five variants).
    C fI[Q];
    C rI = 1, uI[3] = {0, 0, 0}, rB = 1, uB[3] = {0, 0, 0}, dI = 0, dB = 0, u2I = 0, u2B = 0;
    if (ns) {
        const int ii = i - oi, jj = j - oj, kk = k - ok;
        const long long I = (long long)ii + (long long)g.px * ((long long)jj + (long long)g.py * kk);
        if (a.mask[I] & MK_SOLID) {
            int pidx;
            if (axis == 0) pidx = (jj - 1) + ny * (kk - 1);
            else if (axis == 1) pidx = (ii - 1) + nx * (kk - 1);
            else pidx = (ii - 1) + nx * (jj - 1);
            const int slot = a.face_map[face][pidx];
            const C *so = static_cast<const C *>(a.s_old);
#pragma unroll
            for (int q = 0; q < Q; ++q) fI[q] = so[(long long)slot * Q + q];
        } else {
#pragma unroll
            for (int q = 0; q < Q; ++q) fI[q] = read_post<MODE, LAY>(a.pop, a.mask, g, I, q, a.inv_scale);
        }
        const bool plus = (oi + oj + ok) > 0;
        if constexpr (MODE == M_F64) {
            // macro_at (05:545-568)
            double r, jx, jy, jz;
            moments64(fI, r, jx, jy, jz);
            double ux, uy, uz;
            if (r > 1e-30) { ux = jx / r; uy = jy / r; uz = jz / r; }
            else { r = 1.0; ux = uy = uz = 0.0; }
            rI = r; uI[0] = ux; uI[1] = uy; uI[2] = uz;
            const double pI = CS2 * (rI - 1.0);
            double Wp, Wm;
            if (!plus) { Wp = 0.0; Wm = uI[axis] - pI / CS; }
            else { Wm = 0.0; Wp = uI[axis] + pI / CS; }
            const double unB = 0.5 * (Wp + Wm);
            const double pB = 0.5 * CS * (Wp - Wm);
            rB = 1.0 + pB / CS2;
            uB[0] = uI[0]; uB[1] = uI[1]; uB[2] = uI[2];
            uB[axis] = unB;
            u2I = uI[0] * uI[0] + uI[1] * uI[1] + uI[2] * uI[2];
            u2B = uB[0] * uB[0] + uB[1] * uB[1] + uB[2] * uB[2];
        } else {
            float dr, jx, jy, jz;
            moments32(fI, dr, jx, jy, jz);
            float r = 1.0f + dr;
            float ux, uy, uz;
            if (r > 1e-30f) { const float ir = 1.0f / r; ux = jx * ir; uy = jy * ir; uz = jz * ir; }
            else { r = 1.0f; dr = 0.0f; ux = uy = uz = 0.0f; }
            rI = r; dI = dr; uI[0] = ux; uI[1] = uy; uI[2] = uz;
            const float pI = (float)CS2 * dr;
            float Wp, Wm;
            if (!plus) { Wp = 0.0f; Wm = uI[axis] - pI * (float)(1.0 / CS); }
            else { Wm = 0.0f; Wp = uI[axis] + pI * (float)(1.0 / CS); }
            const float unB = 0.5f * (Wp + Wm);
            const float pB = 0.5f * (float)CS * (Wp - Wm);
            dB = 3.0f * pB;
            rB = 1.0f + dB;
            uB[0] = uI[0]; uB[1] = uI[1]; uB[2] = uI[2];
            uB[axis] = unB;
            u2I = 1.5f * (uI[0] * uI[0] + uI[1] * uI[1] + uI[2] * uI[2]);
            u2B = 1.5f * (uB[0] * uB[0] + uB[1] * uB[1] + uB[2] * uB[2]);
        }
    }

#pragma unroll
    for (int q = 1; q < Q; ++q) {
        const int ti = i + EX[q], tj = j + EY[q], tk = k + EZ[q];
        if (ti < 1 || ti > nx || tj < 1 || tj > ny || tk < 1 || tk > nz) continue;
        C val = rest<MODE>(q);
        if (ns) {
            if constexpr (MODE == M_F64) {
                // ghost[q] = feqB[q] + (fI[q] - feqI[q])  (05:705-707)
                const double feI = feq64(q, rI, uI[0], uI[1], uI[2], u2I);
                const double feB = feq64(q, rB, uB[0], uB[1], uB[2], u2B);
                val = feB + (fI[q] - feI);
            } else {
                const float geI = geq32(q, dI, rI, uI[0], uI[1], uI[2], u2I);
                const float geB = geq32(q, dB, rB, uB[0], uB[1], uB[2], u2B);
                val = geB + (fI[q] - geI);
            }
        }
        const long long tc = h + nb_off(g, q);
        long long idx;
        if constexpr (LAY == LAY_P) {
            idx = (long long)INV[q] * N + h;
        } else {
            idx = (a.mask[tc] & MK_SOLID) ? (long long)INV[q] * N + h : (long long)q * N + tc;
        }
        st<MODE>(a.pop, idx, val, a.scale);
    }
}

// ---- solid cells on open faces (05:917-938 for a SOLID dst) -------------------------------
// f_next[b][q] = solid(b - e_q) ? f[b][inv q] : f*_q(b - e_q).  With b solid, f*_q of a
// fluid neighbour sits in slot INV[q] of that neighbour in either layout; the ghost kernel
// puts a halo neighbour's value in the same place.
template <int MODE>
__global__ __launch_bounds__(256) void k_solid(SolidArgs a) {
    using C = typename Tr<MODE>::C;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= a.n) return;
    const long long c = a.cells[n];
    const long long N = a.g.ncell;
    const C *so = static_cast<const C *>(a.s_old);
    C *sn = static_cast<C *>(a.s_new);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const long long src = c - nb_off(a.g, q);
        C v;
        if (a.mask[src] & MK_SOLID) {
            v = so[(long long)n * Q + INV[q]];
            if (a.alpha != 1.0) v = reflect<MODE>(q, v, a.alpha);
        } else {
            v = ld<MODE>(a.pop, (long long)INV[q] * N + src, a.inv_scale);
        }
        sn[(long long)n * Q + q] = v;
    }
}

// ---- source injection (05:660-665) ---------------------------------------------------------
template <int MODE, int LAY>
__global__ __launch_bounds__(256) void k_source(SourceArgs a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n * Q) return;
    const int e = t / Q, q = t % Q;
    const long long c = a.cells[e];
    const long long N = a.g.ncell;
    long long idx;
    if constexpr (LAY == LAY_P) {
        idx = (long long)INV[q] * N + c;
    } else {
        const long long tc = c + nb_off(a.g, q);
        idx = (a.mask[tc] & MK_SOLID) ? (long long)INV[q] * N + c : (long long)q * N + tc;
    }
    if constexpr (MODE == M_F64) {
        double *p = static_cast<double *>(a.pop);
        p[idx] += W64[q] * a.adds[e];
    } else {
        const float v = ld<MODE>(a.pop, idx, a.inv_scale) + (float)(W64[q] * a.adds[e]);
        st<MODE>(a.pop, idx, v, a.scale);
    }
}

// ---- initialisation and gathers ------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(256) void k_fill_rest(void *pop, long long ncell) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell) return;
#pragma unroll
    for (int q = 0; q < Q; ++q) st<MODE>(pop, (long long)q * ncell + t, rest<MODE>(q), 1.0f);
}

template <int MODE>
__global__ __launch_bounds__(256) void k_set_density(void *pop, const uint8_t *mask, GridDesc g,
                                                    const double *rho, float scale) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)g.nx * g.ny * g.nz;
    if (n >= total) return;
    const int i = (int)(n % g.nx) + 1;
    const int j = (int)((n / g.nx) % g.ny) + 1;
    const int k = (int)(n / ((long long)g.nx * g.ny)) + 1;
    const long long c = (long long)i + (long long)g.px * ((long long)j + (long long)g.py * k);
    if (mask[c] & MK_SOLID) return;
    const double r = rho[n];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        // feq(q, rho, 0, 0, 0) = w_q * rho (05:286-291 with u = 0); symmetric in q, so the
        // value is the same in either slot convention.
        if constexpr (MODE == M_F64) st<MODE>(pop, (long long)q * g.ncell + c, W64[q] * r, scale);
        else st<MODE>(pop, (long long)q * g.ncell + c, (float)(W64[q] * (r - 1.0)), scale);
    }
}

template <int MODE>
__global__ __launch_bounds__(256) void k_fill_state_rest(void *state, int n) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * Q) return;
    using C = typename Tr<MODE>::C;
    static_cast<C *>(state)[t] = rest<MODE>((int)(t % Q));
}

template <int MODE, int LAY>
__global__ __launch_bounds__(256) void k_gather_f(const void *pop, const uint8_t *mask, GridDesc g,
                                                 float inv_scale, double *out) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)g.nx * g.ny * g.nz;
    if (n >= total) return;
    const int i = (int)(n % g.nx) + 1;
    const int j = (int)((n / g.nx) % g.ny) + 1;
    const int k = (int)(n / ((long long)g.nx * g.ny)) + 1;
    const long long c = (long long)i + (long long)g.px * ((long long)j + (long long)g.py * k);
    const bool solid = mask[c] & MK_SOLID;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        double v;
        if (solid) v = NAN;
        else v = (double)read_post<MODE, LAY>(pop, mask, g, c, q, inv_scale);
        if constexpr (MODE != M_F64) v += W64[q];
        out[n * Q + q] = v;
    }
}

// ---- launch wrappers ---------------------------------------------------------------------

inline unsigned blocks_for(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <int MODE>
hipError_t Launch<MODE>::step(const StepArgs &a, bool odd, bool capture, dim3 block, hipStream_t s) {
    const dim3 grid(a.nblocks);
    if (odd) {
        if (capture) k_step<MODE, true, true><<<grid, block, 0, s>>>(a);
        else k_step<MODE, true, false><<<grid, block, 0, s>>>(a);
    } else {
        if (capture) k_step<MODE, false, true><<<grid, block, 0, s>>>(a);
        else k_step<MODE, false, false><<<grid, block, 0, s>>>(a);
    }
    return hipGetLastError();
}

template <int MODE>
hipError_t Launch<MODE>::ghost(const GhostArgs &a, int lay, hipStream_t s) {
    if (a.n_halo <= 0) return hipSuccess;
    const unsigned nb = blocks_for(a.n_halo, 256);
    if (lay == LAY_P) k_ghost<MODE, LAY_P><<<nb, 256, 0, s>>>(a);
    else k_ghost<MODE, LAY_Q><<<nb, 256, 0, s>>>(a);
    return hipGetLastError();
}

template <int MODE>
hipError_t Launch<MODE>::solid(const SolidArgs &a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    k_solid<MODE><<<blocks_for(a.n, 256), 256, 0, s>>>(a);
    return hipGetLastError();
}

template <int MODE>
hipError_t Launch<MODE>::source(const SourceArgs &a, int lay, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    const unsigned nb = blocks_for((long long)a.n * Q, 256);
    if (lay == LAY_P) k_source<MODE, LAY_P><<<nb, 256, 0, s>>>(a);
    else k_source<MODE, LAY_Q><<<nb, 256, 0, s>>>(a);
    return hipGetLastError();
}

template <int MODE>
hipError_t Launch<MODE>::fill_rest(void *pop, const GridDesc &g, hipStream_t s) {
    k_fill_rest<MODE><<<blocks_for(g.ncell, 256), 256, 0, s>>>(pop, g.ncell);
    return hipGetLastError();
}

template <int MODE>
hipError_t Launch<MODE>::set_density(void *pop, const uint8_t *mask, const GridDesc &g,
                                     const double *rho, float scale, hipStream_t s) {
    const long long total = (long long)g.nx * g.ny * g.nz;
    k_set_density<MODE><<<blocks_for(total, 256), 256, 0, s>>>(pop, mask, g, rho, scale);
    return hipGetLastError();
}

template <int MODE>
hipError_t Launch<MODE>::fill_state_rest(void *state, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_fill_state_rest<MODE><<<blocks_for((long long)n * Q, 256), 256, 0, s>>>(state, n);
    return hipGetLastError();
}

template <int MODE>
hipError_t Launch<MODE>::gather_f(const void *pop, const uint8_t *mask, const GridDesc &g, int lay,
                                  float inv_scale, double *out, hipStream_t s) {
    const long long total = (long long)g.nx * g.ny * g.nz;
    if (lay == LAY_P) k_gather_f<MODE, LAY_P><<<blocks_for(total, 256), 256, 0, s>>>(pop, mask, g, inv_scale, out);
    else k_gather_f<MODE, LAY_Q><<<blocks_for(total, 256), 256, 0, s>>>(pop, mask, g, inv_scale, out);
    return hipGetLastError();
}

}  // namespace nlos
