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

#include <type_traits>

#include "common.hpp"
#include "launch.hpp"

namespace nlos {

template <int MODE> struct Tr;
template <> struct Tr<M_F64> { using S = double; using C = double; };
template <> struct Tr<M_F32> { using S = float; using C = float; };
template <> struct Tr<M_F16> { using S = _Float16; using C = float; };
template <> struct Tr<M_B16> { using S = short; using C = float; };   // see Mode

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

// ---- B16 block floating point (see Mode) --------------------------------------------------
// Cells are coded in normalised units x_q = g_q / w_q.  x_0 (the rest slot, read only by its
// own cell) is kept in FP32.  The exponent e of a cell is the smallest with
// max_{q>=1} |x_q| < 2^e; the mantissas are x_q * 2^(15 - e) rounded to nearest even and
// saturated to 16 bits, and the stored byte is E = e - 15, so that x_q = m_q * 2^E.
// Array of N = ncell cells: bytes [0, 4N) x_0 (FP32), [4N, 40N) mantissas (slot q >= 1 at
// 16-bit index 2N + mant_index(q, c)), [40N, 41N) exponents.

__device__ __forceinline__ const float *b16_rest(const void *pop) { return static_cast<const float *>(pop); }
__device__ __forceinline__ float *b16_rest(void *pop) { return static_cast<float *>(pop); }
__device__ __forceinline__ const short *b16_mants(const void *pop, long long N) {
    return static_cast<const short *>(pop) + 2 * N;
}
__device__ __forceinline__ short *b16_mants(void *pop, long long N) { return static_cast<short *>(pop) + 2 * N; }
__device__ __forceinline__ const signed char *b16_exps(const void *pop, long long N) {
    return static_cast<const signed char *>(pop) + 40 * N;
}
__device__ __forceinline__ signed char *b16_exps(void *pop, long long N) {
    return static_cast<signed char *>(pop) + 40 * N;
}
__device__ __forceinline__ float b16_dec(int m, int E) { return __builtin_amdgcn_ldexpf((float)m, E); }

// Exponent e of a cell (clamped below so that e - 15 fits a signed byte).
__device__ __forceinline__ int b16_exp_of(const float *x) {
    float mx = 0.0f;
#pragma unroll
    for (int q = 1; q < Q; ++q) mx = fmaxf(mx, fabsf(x[q]));
    return max(__builtin_amdgcn_frexp_expf(mx), B16_EMIN);   // mx = frac * 2^e, frac in [0.5, 1)
}

// Mantissas of slots 1..18 at exponent e, two per saturating pack (|x| < 2^e rounds to at
// most 32768, which saturates to 32767).
__device__ __forceinline__ void b16_code(const float *x, int e, short *m) {
    const float sc = __builtin_amdgcn_ldexpf(1.0f, B16_BITS - e);
    m[0] = 0;
#pragma unroll
    for (int q = 1; q < Q; q += 2) {
        const int a = (int)__builtin_rintf(x[q] * sc);
        const int b = q + 1 < Q ? (int)__builtin_rintf(x[q + 1] * sc) : 0;
        const auto p = __builtin_amdgcn_cvt_pk_i16(a, b);
        m[q] = p.x;
        if (q + 1 < Q) m[q + 1] = p.y;
    }
}

// Whole-cell access in deviation units g (the kernels other than the sweep).
__device__ __forceinline__ void b16_load_g(const void *pop, long long N, long long c, float *g) {
    const short *m = b16_mants(pop, N);
    const int E = b16_exps(pop, N)[c];
    g[0] = b16_rest(pop)[c] * wq<float>(0);
#pragma unroll
    for (int q = 1; q < Q; ++q) g[q] = b16_dec(m[mant_index(q, c)], E) * wq<float>(q);
}
__device__ __forceinline__ float b16_slot_g(const void *pop, long long N, long long c, int q) {
    if (q == 0) return b16_rest(pop)[c] * wq<float>(0);
    return b16_dec(b16_mants(pop, N)[mant_index(q, c)], b16_exps(pop, N)[c]) * wq<float>(q);
}
// Codes x into cell c; slots with write[q] false are left as they are (the ghost layer
// writes only the slots that interior cells pull).  Returns e.
__device__ __forceinline__ int b16_store_x(void *pop, long long N, long long c, const float *x,
                                           unsigned write = 0xffffffffu) {
    const int e = b16_exp_of(x);
    short m[Q];
    b16_code(x, e, m);
    if (write & 1u) b16_rest(pop)[c] = x[0];
    short *mp = b16_mants(pop, N);
#pragma unroll
    for (int q = 1; q < Q; ++q)
        if ((write >> q) & 1u) mp[mant_index(q, c)] = m[q];
    b16_exps(pop, N)[c] = (signed char)(e - B16_BITS);
    return e;
}
// g -> x: 1/w_q = 3, 18, 36 are exact in FP32.
__device__ __forceinline__ void b16_g_to_x(const float *g, float *x) {
#pragma unroll
    for (int q = 0; q < Q; ++q) x[q] = g[q] * (q == 0 ? 3.0f : (q < 7 ? 18.0f : 36.0f));
}

// Neighbour offset of direction q.
__device__ __forceinline__ long long nb_off(const GridDesc &g, int q) {
    return (long long)EX[q] + (long long)EY[q] * g.px + (long long)EZ[q] * g.plane;
}
// Its y/z part, a multiple of 64 cells.
__device__ __forceinline__ long long yz_off(const GridDesc &g, int q) {
    return (long long)EY[q] * g.px + (long long)EZ[q] * g.plane;
}

// Slot addressing of a sweep lane.  A sweep tile's corner cell cb is a multiple of 64
// (interior rows start at i = 1 with origin 63, tile widths are multiples of 64), so the
// index of slot s of cell cb + u is b0 + s*ss + o(u) with b0 uniform over the workgroup and
// o(u) >= 0 for u >= -64:
//   blocked, NS slots per block: b0 = ((cb >> 6) - 1) * NS*64, ss = 64,
//     o(u) = ((u + 64) >> 6) * NS*64 + (u & 63); a y/z offset d moves an index by d*NS;
//   slot-major: b0 = cb - 64, ss = ncell, o(u) = u + 64; d moves an index by d.
// o[1 + e] holds the lane's o(t - e) for e = -1, 0, 1: the cell itself and its two x
// neighbours.
template <int MODE, int NS = Q>
struct Lanes {
    static constexpr bool BLK = blocked_mode(MODE);
    static constexpr long long YM = BLK ? NS : 1;
    long long b0, ss;
    unsigned o[3];
    __device__ __forceinline__ Lanes(long long cb, unsigned t, long long ncell)
        : b0(BLK ? ((cb >> 6) - 1) * (NS * PBLK) : cb - PBLK), ss(BLK ? PBLK : ncell) {
#pragma unroll
        for (int e = -1; e <= 1; ++e) {
            const unsigned u = t - (unsigned)e + PBLK;   // t - e + 64 >= 63
            o[1 + e] = BLK ? (u >> 6) * (NS * PBLK) + (u & (PBLK - 1)) : u;
        }
    }
    // slot s of the lane's cell c
    __device__ __forceinline__ long long self(int s) const { return b0 + s * ss + o[1]; }
    // slot s of c - e_q (pull source)
    __device__ __forceinline__ long long from(const GridDesc &g, int q, int s) const {
        return b0 - yz_off(g, q) * YM + s * ss + o[1 + EX[q]];
    }
    // slot s of c + e_q (push target)
    __device__ __forceinline__ long long to(const GridDesc &g, int q, int s) const {
        return b0 + yz_off(g, q) * YM + s * ss + o[1 - EX[q]];
    }
};

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
// ((1-s)f + s*w - w = (1-s)g).  Works on the stored values G = g*S (S = 1 for FP32, 2^k
// for FP16): the moments are scaled back once and the equilibrium is formed in stored
// units.  Ordering B pairs every direction with its inverse; a pair shares the even part
// of the equilibrium (cu_inv = -cu), and its sum and difference give the moments.
//   g_post = keep*((1-omega)*g + omega*geq),  geq = w*(drho + rho*(3cu + 4.5cu^2 - 1.5u^2))
// Returns drho.
__device__ __forceinline__ float collide32(float *G, float omega, float keep, float S, float invS) {
    constexpr int PAIR[9] = {1, 3, 5, 7, 8, 11, 12, 15, 16};
    float sp[9], dp[9];
#pragma unroll
    for (int n = 0; n < 9; ++n) {
        sp[n] = G[PAIR[n]] + G[INV[PAIR[n]]];
        dp[n] = G[PAIR[n]] - G[INV[PAIR[n]]];
    }
    // pairs: 1 (1,0,0) 3 (0,1,0) 5 (0,0,1) 7 (1,1,0) 8 (1,-1,0) 11 (1,0,1) 12 (1,0,-1)
    //        15 (0,1,1) 16 (0,1,-1)
    float drs = G[0] + (((sp[0] + sp[1]) + (sp[2] + sp[3])) + ((sp[4] + sp[5]) + (sp[6] + sp[7])) + sp[8]);
    const float jx = ((dp[0] + dp[3]) + (dp[4] + dp[5])) + dp[6];
    const float jy = ((dp[1] + dp[3]) + (dp[7] + dp[8])) - dp[4];
    const float jz = ((dp[2] + dp[5]) + (dp[7] - dp[6])) - dp[8];
    float dr = drs * invS;
    float r = 1.0f + dr;
    if (r <= 0.0f) { r = 1e-12f; dr = r - 1.0f; drs = dr * S; }
    const float k = invS * __builtin_amdgcn_rcpf(r);
    const float ux = jx * k, uy = jy * k, uz = jz * k;
    const float u2c = 1.5f * (ux * ux + uy * uy + uz * uz);
    const float rS = r * S;
    const float a = keep * (1.0f - omega);      // weight of the old population
    const float bw = keep * omega;               // weight of the equilibrium
    G[0] = a * G[0] + (bw * wq<float>(0)) * (drs - rS * u2c);
    // per weight class (1/18: pairs 0-2, 1/36: pairs 3-8)
    const float c1a = bw * wq<float>(1) * drs, c2a = bw * wq<float>(1) * rS, c3a = 3.0f * c2a;
    const float c1b = bw * wq<float>(7) * drs, c2b = bw * wq<float>(7) * rS, c3b = 3.0f * c2b;
#pragma unroll
    for (int n = 0; n < 9; ++n) {
        const int q = PAIR[n], p = INV[q];
        const float c1 = n < 3 ? c1a : c1b, c2 = n < 3 ? c2a : c2b, c3 = n < 3 ? c3a : c3b;
        const float cu = cu_of<float>(q, ux, uy, uz);
        const float x = (4.5f * cu) * cu - u2c;
        const float even = c2 * x + c1;
        const float odd = c3 * cu;
        G[q] = a * G[q] + (even + odd);
        G[p] = a * G[p] + (even - odd);
    }
    return dr;
}

// BGK collision of collide32 on normalised deviations x_q = g_q / w_q (B16 storage): the
// weights move into the moment sums and drop out of the update,
//   x*_q = keep*((1-omega)*x_q + omega*(drho + rho*(3cu + 4.5cu^2 - 1.5u^2))).
// Returns drho.
__device__ __forceinline__ float collide_n(float *X, float omega, float keep) {
    constexpr int PAIR[9] = {1, 3, 5, 7, 8, 11, 12, 15, 16};
    float sp[9], dp[9];
#pragma unroll
    for (int n = 0; n < 9; ++n) {
        sp[n] = X[PAIR[n]] + X[INV[PAIR[n]]];
        dp[n] = X[PAIR[n]] - X[INV[PAIR[n]]];
    }
    constexpr float W0 = (float)(1.0 / 3.0), W1 = (float)(1.0 / 18.0), W2 = (float)(1.0 / 36.0);
    float dr = X[0] * W0 + (sp[0] + sp[1] + sp[2]) * W1 +
               (((sp[3] + sp[4]) + (sp[5] + sp[6])) + (sp[7] + sp[8])) * W2;
    const float jx = dp[0] * W1 + (((dp[3] + dp[4]) + dp[5]) + dp[6]) * W2;
    const float jy = dp[1] * W1 + ((dp[3] - dp[4]) + (dp[7] + dp[8])) * W2;
    const float jz = dp[2] * W1 + ((dp[5] - dp[6]) + (dp[7] - dp[8])) * W2;
    float r = 1.0f + dr;
    if (r <= 0.0f) { r = 1e-12f; dr = r - 1.0f; }
    const float ir = __builtin_amdgcn_rcpf(r);
    const float ux = jx * ir, uy = jy * ir, uz = jz * ir;
    const float u2c = 1.5f * (ux * ux + uy * uy + uz * uz);
    const float a = keep * (1.0f - omega);
    const float bw = keep * omega;
    const float c1 = bw * dr, c2 = bw * r, c3 = 3.0f * c2;
    X[0] = a * X[0] + (c1 - c2 * u2c);
#pragma unroll
    for (int n = 0; n < 9; ++n) {
        const int q = PAIR[n], p = INV[q];
        const float cu = cu_of<float>(q, ux, uy, uz);
        const float even = c2 * ((4.5f * cu) * cu - u2c) + c1;
        const float odd = c3 * cu;
        X[q] = a * X[q] + (even + odd);
        X[p] = a * X[p] + (even - odd);
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

// Lane access base[t]: the tile base is the same for every lane of the wave, the lane
// offset t is 32-bit.
template <int MODE>
__device__ __forceinline__ typename Tr<MODE>::C ldu(const typename Tr<MODE>::S *base, unsigned t, float inv_scale) {
    using S = typename Tr<MODE>::S;
    const S v = base[t];
    if constexpr (MODE == M_F16) return static_cast<float>(v) * inv_scale;
    else return v;
}

template <int MODE>
__device__ __forceinline__ void stu(typename Tr<MODE>::S *base, unsigned t, typename Tr<MODE>::C v, float scale) {
    using S = typename Tr<MODE>::S;
    if constexpr (MODE == M_F16) base[t] = static_cast<S>(v * scale);
    else base[t] = v;
}

// Whole-wave shift by one lane: with right = true lane L receives lane L-1's value and lane
// 0 receives `edge`; with right = false lane L receives lane L+1's value and lane 63 `edge`.
// Needs every lane of the wave active.
__device__ __forceinline__ int dpp_shift(int v, int edge, bool right) {
    return right ? __builtin_amdgcn_update_dpp(edge, v, 0x138 /* wave_shr:1 */, 0xf, 0xf, false)
                 : __builtin_amdgcn_update_dpp(edge, v, 0x130 /* wave_shl:1 */, 0xf, 0xf, false);
}

template <class S>
__device__ __forceinline__ S wave_shift(S v, S edge, bool right) {
    if constexpr (sizeof(S) == 2) {
        const int r = dpp_shift((int)__builtin_bit_cast(unsigned short, v),
                                (int)__builtin_bit_cast(unsigned short, edge), right);
        return __builtin_bit_cast(S, (unsigned short)r);
    } else if constexpr (sizeof(S) == 4) {
        return __builtin_bit_cast(S, dpp_shift(__builtin_bit_cast(int, v), __builtin_bit_cast(int, edge), right));
    } else {
        const long long bv = __builtin_bit_cast(long long, v), be = __builtin_bit_cast(long long, edge);
        const int lo = dpp_shift((int)bv, (int)be, right);
        const int hi = dpp_shift((int)(bv >> 32), (int)(be >> 32), right);
        return __builtin_bit_cast(S, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
    }
}

template <int MODE, bool ODD, bool CAPTURE>
__global__ __launch_bounds__(256) void k_step(StepArgs a) {
    using C = typename Tr<MODE>::C;
    using S = typename Tr<MODE>::S;
    // Values stay in stored units through the sweep (scaled FP16 is converted, not rescaled).
    constexpr float RAW = 1.0f;
    const int b = blockIdx.x;
    const int bi = b % a.tiles_x;
    const int rest_b = b / a.tiles_x;
    const int bj = rest_b % a.tiles_y;
    const int k = rest_b / a.tiles_y + 1;
    const int i0 = bi * (int)blockDim.x + 1;
    const int j0 = bj * (int)blockDim.y + 1;
    const int i = i0 + (int)threadIdx.x;
    const int j = j0 + (int)threadIdx.y;
    if (i > a.g.nx || j > a.g.ny) return;

    const long long N = a.g.ncell;
    const long long cb = cell_index(a.g, i0, j0, k);                  // block corner, uniform
    const unsigned t = threadIdx.x + (unsigned)a.g.px * threadIdx.y;  // lane offset
    const long long c = cb + t;
    const Lanes<MODE> L(cb, t, N);
    S *P = static_cast<S *>(a.pop);
    const C *sx = static_cast<const C *>(a.sig_x);
    const C *sy = static_cast<const C *>(a.sig_y);
    const C *sz = static_cast<const C *>(a.sig_z);

    // Every load is issued before the mask is looked at, so that the mask byte adds no
    // memory round trip in front of the population loads; all addresses are inside the
    // lattice for any interior cell (solid cells included).
    //   ODD (LAY_P): f_in_q(c) = f*_q(c - e_q) from slot INV[q] of c - e_q.  When c - e_q is
    //     solid, that slot of the solid cell holds the bounced f*_{INV q}(c): the previous
    //     even sweep (or k_prime_walls) put it there.
    //   EVEN (LAY_Q): streamed and bounced populations both sit in slot q of c.
    const uint8_t m = (a.mask + cb)[t];
    C f[Q];
    if (ODD && a.dpp) {
        // Odd sweep with every lane active: a wave is one aligned row of 64 cells, so a
        // source row shifted by one cell in x is read as its aligned row plus one edge
        // element and moved across lanes, instead of as a load straddling two lines.
        // the edge elements: the cell one before the row start, or one past its end
        const long long row = cb + (long long)a.g.px * (long long)threadIdx.y;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const long long rowoff = yz_off(a.g, q);
            const S *base = P + (L.b0 - rowoff * L.YM + INV[q] * L.ss);   // + o[1]: slot INV[q] of c - rowoff
            if (EX[q] == 0) {
                f[q] = ldu<MODE>(base, L.o[1], RAW);
            } else {
                const S v = base[L.o[1]];
                const S edge = P[pop_index<MODE>(INV[q], (EX[q] > 0 ? row - 1 : row + 64) - rowoff, N)];
                f[q] = (C)wave_shift<S>(v, edge, EX[q] > 0);
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const long long idx = ODD ? L.from(a.g, q, INV[q]) : L.self(q);
            f[q] = ldu<MODE>(P + idx, 0u, RAW);
        }
    }
    // sigma(min d) = max over axes of sigma(d_axis): the ramp is monotone (05:491-506).
    const C sig = fmax(sx[i], fmax(sy[j], sz[k]));
    if (m & MK_SOLID) return;
    const bool near = m & MK_NEAR;

    uint32_t sb = 0;   // bit q: c + e_q is solid (only gathered for cells next to a wall)
    if (near) {
#pragma unroll
        for (int q = 1; q < Q; ++q)
            if (a.mask[c + nb_off(a.g, q)] & MK_SOLID) sb |= 1u << q;
        // EVEN: a population that came off a wall is reflected here (ODD: the reflected
        // value was stored in the wall's slot).
        if (!ODD && a.alpha != 1.0) {
#pragma unroll
            for (int q = 1; q < Q; ++q)
                if ((sb >> INV[q]) & 1u) f[q] = reflect<MODE>(q, f[q], a.alpha);
        }
    }

    double dval;   // rho - 1
    bool finite;
    if constexpr (MODE == M_F64) {
        const double r = collide64(f, a.omega, sig);
        dval = r - 1.0;
        finite = isfinite(r);
    } else {
        const float dr = collide32(f, (float)a.omega, 1.0f - sig, a.scale, a.inv_scale);
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
    if (i == 1 || i == a.g.nx) {
#pragma unroll
        for (int face = 0; face < 2; ++face) {
            if ((face == 0 ? i != 1 : i != a.g.nx) || !a.fb[face]) continue;
            S *fb = static_cast<S *>(a.fb[face]) + ((long long)(j - 1) + (long long)a.g.ny * (k - 1)) * Q;
#pragma unroll
            for (int q = 0; q < Q; ++q) st<MODE>(fb, q, f[q], RAW);   // per-lane address
        }
    }

    if constexpr (ODD) {
        // -> LAY_Q: slot q of c + e_q; when c + e_q is solid, slot INV[q] of c (half-way
        // bounce-back folded into the store; the copy into the wall's slot is never read).
#pragma unroll
        for (int q = 0; q < Q; ++q) stu<MODE>(P + L.to(a.g, q, q), 0u, f[q], RAW);
        if (sb) {
#pragma unroll
            for (int q = 1; q < Q; ++q)
                if ((sb >> q) & 1u) st<MODE>(a.pop, L.self(INV[q]), f[q], RAW);
        }
    } else {
        // -> LAY_P: slot INV[q] of c; a population headed into a wall is also stored,
        // reflected, in slot q of the wall cell c + e_q, where the next odd sweep reads it.
#pragma unroll
        for (int q = 0; q < Q; ++q) stu<MODE>(P + L.self(INV[q]), 0u, f[q], RAW);
        if (sb) {
#pragma unroll
            for (int q = 1; q < Q; ++q)
                if ((sb >> q) & 1u)
                    st<MODE>(a.pop, L.to(a.g, q, q),
                             a.alpha != 1.0 ? reflect<MODE>(q, f[q], a.alpha) : f[q], RAW);
        }
    }
}

// Pull sweep of the two-array pattern (LAY_AB, 05:917-968): f_in_q(c) = cur[q][c - e_q],
// or the half-way bounce-back cur[INV q][c] when c - e_q is solid; collide; write
// nxt[q][c] (aligned, local).  Solid lanes store the rest state into their own slots,
// which nothing reads: every wave then writes whole 128-byte lines, where a line with
// unwritten lanes costs the memory a read-modify-write.  One thread per cell, one plane
// per workgroup tile, per-lane addressing through Lanes (a workgroup-uniform base plus a
// 32-bit lane offset; the compiler keeps every slot offset as a constant).  Pulls from
// the x-halo cells come from the x-halo records when StepArgs::xh is set.

template <class S>
__device__ __forceinline__ S read_lane(S v, int src) {
    if constexpr (sizeof(S) == 2) {
        const int b = __builtin_bit_cast(unsigned short, v);
        return __builtin_bit_cast(S, (unsigned short)__builtin_amdgcn_readlane(b, src));
    } else if constexpr (sizeof(S) == 4) {
        return __builtin_bit_cast(S, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), src));
    } else {
        const long long b = __builtin_bit_cast(long long, v);
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, src);
        const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), src);
        return __builtin_bit_cast(S, (long long)(((unsigned long long)hi << 32) | lo));
    }
}

// rec[q] = f[q] of lane src, written by lanes 0..Q-1 with one store instruction (a single
// lane storing Q scattered values costs the sweep far more).  Wave-uniform call; lanes
// 0..Q-1 must be active.
template <class S, class C>
__device__ __forceinline__ void store_record(S *rec, const C *f, int src, int lane) {
    S mine = S(0);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const S r = read_lane(static_cast<S>(f[q]), src);
        if (lane == q) mine = r;
    }
    if (lane < Q) rec[lane] = mine;
}

// The same pointer in the global address space.  Arguments read from memory (the
// multi-scene sweep) are generic pointers, which would compile to flat accesses.
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T *glb(T *p) {
    return (__attribute__((address_space(1))) T *)p;
}

// Stored populations and mask byte of one cell (FP16 values two to a register).
template <int MODE>
struct AbIn {
    typename Tr<MODE>::S v[Q];
    uint8_t m;
    __device__ __forceinline__ typename Tr<MODE>::C get(int q) const { return static_cast<typename Tr<MODE>::C>(v[q]); }
};

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

template <>
struct AbIn<M_F16> {
    half2_t v[(Q + 1) / 2];
    uint8_t m;
    __device__ __forceinline__ float get(int q) const {
        return static_cast<float>((q & 1) ? v[q >> 1].y : v[q >> 1].x);
    }
};

// B16: the cell's own x_0, and for q >= 1 the mantissa of slot q of x - e_q and that cell's
// exponent (e[0] is the cell's own); get() is the normalised value x_q.
template <>
struct AbIn<M_B16> {
    float r0;
    short v[Q];
    signed char e[Q];
    uint8_t m;
    __device__ __forceinline__ float get(int q) const { return q == 0 ? r0 : b16_dec(v[q], e[q]); }
};

template <int MODE, bool CAPTURE, bool VEL = false>
__device__ __forceinline__ void ab_cell(const StepArgs &a, const AbIn<MODE> &in, typename Tr<MODE>::C sig,
                                        int i, int j, int k, long long cb, unsigned t) {
    using C = typename Tr<MODE>::C;
    using S = typename Tr<MODE>::S;
    const long long N = a.g.ncell;
    const long long c = cb + t;
    // lattice slot s0 + r holds slot r; the B16 mantissa arrays hold slots 1..18
    constexpr int S0 = MODE == M_B16 ? 1 : 0;
    const std::conditional_t<MODE == M_B16, Lanes<M_B16, Q - 1>, Lanes<MODE>> L(cb, t, N);
    const uint8_t m = in.m;
    C f[Q];
    if (m & MK_SOLID) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            if constexpr (MODE == M_F64) f[q] = W64[q];
            else f[q] = 0.0f;
        }
    } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) f[q] = in.get(q);
        if (m & MK_NEAR) {
            // all bounce candidates are loaded together with the wall bits (one memory
            // round trip), then selected; loads guarded one by one would serialise
            const uint32_t sb = glb(a.wall_bits)[c];   // bit q: c - e_q is solid
            const auto cur = glb(static_cast<const S *>(MODE == M_B16 ? (const void *)b16_mants(a.pop, N) : a.pop));
            C bb[Q];
#pragma unroll
            for (int q = 1; q < Q; ++q) {
                // B16: slot INV[q] of the cell itself, in its own exponent (w_INV[q] = w_q)
                if constexpr (MODE == M_B16) bb[q] = b16_dec(cur[L.self(INV[q] - S0)], in.e[0]);
                else bb[q] = static_cast<C>(cur[L.self(INV[q])]);
            }
#pragma unroll
            for (int q = 1; q < Q; ++q) {
                if ((sb >> q) & 1u) f[q] = a.alpha != 1.0 ? reflect<MODE>(q, bb[q], a.alpha) : bb[q];
            }
        }

        if constexpr (VEL) {
            // u = j / rho of the pulled populations, which is all the transposed collision
            // needs (k_adjoint): recorded here so the adjoint does not pull the state again
            C r = C(0), jx = C(0), jy = C(0), jz = C(0);
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                r += f[q];
                jx += f[q] * (C)EX[q];
                jy += f[q] * (C)EY[q];
                jz += f[q] * (C)EZ[q];
            }
            if constexpr (MODE != M_F64) r += C(1);   // deviations: rho = 1 + sum g
            C *u = static_cast<C *>(a.vel);
            u[c] = jx / r;
            u[N + c] = jy / r;
            u[2 * N + c] = jz / r;
        }
        double dval;   // rho - 1
        bool finite;
        if constexpr (MODE == M_F64) {
            const double r = collide64(f, a.omega, sig);
            dval = r - 1.0;
            finite = isfinite(r);
        } else if constexpr (MODE == M_B16) {
            const float dr = collide_n(f, (float)a.omega, 1.0f - sig);
            dval = dr;
            finite = isfinite(dr);
        } else {
            const float dr = collide32(f, (float)a.omega, 1.0f - sig, a.scale, a.inv_scale);
            dval = dr;
            finite = isfinite(dr);
        }
        if (!finite) atomicOr(a.flag, 1);
        if (m & MK_PROBE) {
            const int u = probe_slot(a.probe_cells, a.n_probe_cells, c);
            double *row = a.step_ctr ? a.hist + (*a.step_ctr - a.ctr_lag) * a.n_probe_cells : a.probe_row;
            if (u >= 0) row[u] = dval;
        }
        if constexpr (CAPTURE) {
            a.field[(long long)(i - 1) + (long long)a.g.nx * ((long long)(j - 1) + (long long)a.g.ny * (k - 1))] =
                (float)dval;
        }
    }
    // B16: code the cell first, so that the x-face records carry the stored values
    [[maybe_unused]] int be = 0;
    [[maybe_unused]] short bm[Q];
    if constexpr (MODE == M_B16) {
        be = b16_exp_of(f);
        b16_code(f, be, bm);
    }
    // x-face records of the cells (1, j, k) and (nx, j, k) when fluid (see StepArgs::fb);
    // every test below is uniform over the wave (one row segment of 64 cells).
    const int lane = (int)(threadIdx.x & 63u);
    const int iw = i - lane;   // first cell of this wave
#pragma unroll
    for (int face = 0; face < 2; ++face) {
        if (!a.fb[face]) continue;
        if (face == 0 ? iw != 1 : !(iw <= a.g.nx && a.g.nx < iw + 64)) continue;
        const int src = face == 0 ? 0 : a.g.nx - iw;
        if (__builtin_amdgcn_readlane((int)m, src) & MK_SOLID) continue;
        // B16 records hold the stored values decoded to FP32 deviations g (see StepArgs::fb)
        using R = std::conditional_t<MODE == M_B16, float, S>;
        C fr[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            if constexpr (MODE == M_B16) fr[q] = (q == 0 ? f[0] : b16_dec(bm[q], be - B16_BITS)) * wq<float>(q);
            else fr[q] = f[q];
        }
        R *rec = static_cast<R *>(a.fb[face]) + ((long long)(j - 1) + (long long)a.g.ny * (k - 1)) * Q;
        if (a.g.nx - iw >= Q - 1) {
            store_record(rec, fr, src, lane);
        } else if (lane == src) {   // fewer than Q active lanes in this row segment
#pragma unroll
            for (int q = 0; q < Q; ++q) rec[q] = static_cast<R>(fr[q]);
        }
    }
    // Non-temporal stores: the written lines are not read again this step.  Measured at
    // 1024x1024x256: B16 4.95 -> 4.73 ms, F16S 4.19 -> 4.07, F32 8.18 -> 7.85, F64 14.10 ->
    // 13.70, with the L2 fetch volume unchanged (the gain is on the write path).  The same
    // hint on the loads was slower.
    const auto nxt = glb(static_cast<S *>(a.pop_out));
    if constexpr (MODE == M_B16) {
        __builtin_nontemporal_store(f[0], glb(b16_rest(a.pop_out)) + cb + t);
        const auto nm = glb(b16_mants(a.pop_out, N));
#pragma unroll
        for (int q = 1; q < Q; ++q) __builtin_nontemporal_store(bm[q], nm + L.self(q - S0));
        (glb(b16_exps(a.pop_out, N)) + cb)[t] = (signed char)(be - B16_BITS);
    } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) __builtin_nontemporal_store(static_cast<S>(f[q]), nxt + L.self(q));
    }
}

// Tile b: (bi, bj) of the x-y plane at plane k, tw x th cells; the thread sweeps cell
// (tx, ty) of it.  tw is a multiple of 64 and the threads of a wave are consecutive in tx,
// so a wave is one 64-cell row segment.
template <int MODE, bool CAPTURE, bool VEL = false>
__device__ __forceinline__ void step_ab_body(const StepArgs &a, int b, int tx, int ty, int tw, int th) {
    using C = typename Tr<MODE>::C;
    using S = typename Tr<MODE>::S;
    const int bi = b % a.tiles_x;
    const int rest_b = b / a.tiles_x;
    const int bj = rest_b % a.tiles_y;
    const int k = rest_b / a.tiles_y + 1;
    const int i0 = bi * tw + 1;
    const int j0 = bj * th + 1;
    const int i = i0 + tx;
    const int j = j0 + ty;
    if (i > a.g.nx || j > a.g.ny) return;
    const long long N = a.g.ncell;
    // cell = tile corner (the same for the whole workgroup) + lane offset (32-bit): the
    // compiler then forms every slot address as a scalar base plus a 32-bit lane offset
    const long long cb = cell_index(a.g, i0, j0, k);
    const unsigned t = (unsigned)tx + (unsigned)a.g.px * (unsigned)ty;
    const long long c = cb + t;
    constexpr int S0 = MODE == M_B16 ? 1 : 0;   // see ab_cell
    const std::conditional_t<MODE == M_B16, Lanes<M_B16, Q - 1>, Lanes<MODE>> L(cb, t, N);
    AbIn<MODE> in;
    // x-halo sources (see StepArgs::xh): slot q of (i - EX[q], j', k') with i - EX[q] = 0 or
    // nx + 1 comes from the record of row (j', k') instead of the lattice
    const bool xlo = a.xh && i == 1, xhi = a.xh && i == a.g.nx;
    const auto xsrc = [&](int q) { return EX[q] != 0 && (EX[q] > 0 ? xlo : xhi); };
    if constexpr (MODE == M_B16) {
        const auto ex = glb(b16_exps(a.pop, N));
        const auto mn = glb(b16_mants(a.pop, N));
        const auto xm = glb(static_cast<const short *>(a.xh));
        const auto xe = glb(static_cast<const signed char *>(a.xh));
        in.r0 = (glb(b16_rest(a.pop)) + cb)[t];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            if (xsrc(q)) in.e[q] = xe[xh_exp_offset(a.g, EX[q] > 0 ? 0 : 1, j - EY[q], k - EZ[q])];
            else in.e[q] = (ex + (cb - nb_off(a.g, q)))[t];
        }
#pragma unroll
        for (int q = 1; q < Q; ++q) {
            if (xsrc(q)) in.v[q] = xm[xh_index(a.g, q, j - EY[q], k - EZ[q])];
            else in.v[q] = mn[L.from(a.g, q, q - S0)];
        }
        in.v[0] = 0;
    } else {
        const auto cur = glb(static_cast<const S *>(a.pop));
        const auto xs = glb(static_cast<const S *>(a.xh));
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            S v;
            if (xsrc(q)) v = xs[xh_index(a.g, q, j - EY[q], k - EZ[q])];
            else v = cur[L.from(a.g, q, q)];
            if constexpr (MODE == M_F16) {
                if (q & 1) in.v[q >> 1].y = v;
                else in.v[q >> 1].x = v;
            } else {
                in.v[q] = v;
            }
        }
        if constexpr (MODE == M_F16) in.v[Q >> 1].y = _Float16(0);
    }
    in.m = glb(a.mask)[c];
    // sigma(min d) = max over axes of sigma(d_axis): the ramp is monotone (05:491-506).
    const C sig = fmax(glb(static_cast<const C *>(a.sig_x))[i],
                       fmax(glb(static_cast<const C *>(a.sig_y))[j], glb(static_cast<const C *>(a.sig_z))[k]));
    ab_cell<MODE, CAPTURE, VEL>(a, in, sig, i, j, k, cb, t);
}

template <int MODE, bool CAPTURE, bool VEL = false>
__global__ __launch_bounds__(256) void k_step_ab(StepArgs a) {
    step_ab_body<MODE, CAPTURE, VEL>(a, blockIdx.x, (int)threadIdx.x, (int)threadIdx.y, (int)blockDim.x,
                                     (int)blockDim.y);
}

// Sweep over the tiles of several scenes: workgroup (t, s) sweeps tile t of scene s
// (blockIdx.y names the scene; workgroups past a scene's tile count return at once).
template <int MODE>
__global__ __launch_bounds__(256) void k_step_ab_multi(const StepArgs *__restrict__ scenes,
                                                       const int *__restrict__ n_tiles) {
    const int sc = blockIdx.y;
    if ((int)blockIdx.x >= n_tiles[sc]) return;
    // a private copy: read through the pointer, every field would be reloaded after each
    // store of the sweep (the compiler cannot rule out aliasing)
    const StepArgs a = scenes[sc];
    // one-dimensional workgroups of 256 threads; each scene's own tile shape, tile_w x
    // (256 / tile_w) cells with tile_w = 64, 128 or 256
    const int lw = 31 - __builtin_clz((unsigned)a.tile_w);
    step_ab_body<MODE, false>(a, blockIdx.x, (int)(threadIdx.x & ((unsigned)a.tile_w - 1u)), (int)(threadIdx.x >> lw),
                              a.tile_w, 256 >> lw);
}

// The persistent form of the multi-scene sweep: a grid sized to the device, whose
// workgroups walk the (scene, tile) work list, workgroup b taking items b, b + gridDim.x, ...
// (no workgroup of an empty tile, and one launch shape for any set of scenes).
template <int MODE>
__global__ __launch_bounds__(256) void k_step_ab_walk(const StepArgs *__restrict__ scenes,
                                                      const int2 *__restrict__ items, int n_items) {
    for (int it = (int)blockIdx.x; it < n_items; it += (int)gridDim.x) {
        const int2 w = items[it];
        const StepArgs a = scenes[w.x];
        const int lw = 31 - __builtin_clz((unsigned)a.tile_w);
        step_ab_body<MODE, false>(a, w.y, (int)(threadIdx.x & ((unsigned)a.tile_w - 1u)), (int)(threadIdx.x >> lw),
                                  a.tile_w, 256 >> lw);
    }
}

// Wall copies of LAY_P (see Layout): for every fluid cell c next to a wall and every q
// with c + e_q solid, the wall cell receives the reflected f*_q(c) (held in slot INV[q] of
// c) in its slot q.  Needed after the state was set from outside.
template <int MODE>
__global__ __launch_bounds__(256) void k_prime_walls(void *pop, const uint8_t *mask, GridDesc g, double alpha) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)g.nx * g.ny * g.nz;
    if (n >= total) return;
    const int i = (int)(n % g.nx) + 1;
    const int j = (int)((n / g.nx) % g.ny) + 1;
    const int k = (int)(n / ((long long)g.nx * g.ny)) + 1;
    const long long c = cell_index(g, i, j, k);
    const uint8_t m = mask[c];
    if ((m & MK_SOLID) || !(m & MK_NEAR)) return;
    const long long N = g.ncell;
    for (int q = 1; q < Q; ++q) {
        const long long t = c + nb_off(g, q);
        if (!(mask[t] & MK_SOLID)) continue;
        auto v = ld<MODE>(pop, pop_index<MODE>(INV[q], c, N), 1.0f);
        if (alpha != 1.0) v = reflect<MODE>(q, v, alpha);
        st<MODE>(pop, pop_index<MODE>(q, t, N), v, 1.0f);
    }
}

// ---- discrete adjoint of the two-array step -----------------------------------------------
// Forward, per fluid cell x: f_q = S'_q(x - e_q), or the reflected S'_{INV q}(x) when x - e_q
// is solid; f* = (1 - sigma) [f - omega (f - feq(rho, u))] + sigma w; S_{n+1,q}(x) = f*_q;
// probe rho - 1 = sum f.  The transpose, per fluid cell x:
//   mu_p = (1 - sigma) [(1 - omega) lam_p + omega (A (1 + 1.5 u^2) - 4.5 D
//          + e_p . (3 B + 9 C - 3 A u))] (+ dL/dp at probe cells)
// with A = sum w_q lam_q, B = sum w_q lam_q e_q, C = sum w_q lam_q (e_q . u) e_q,
// D = sum w_q lam_q (e_q . u)^2 (the derivative of feq_q = w_q (rho + 3 e.j + 4.5 (e.j)^2/rho
// - 1.5 j^2/rho) with respect to f_p), and mu_q goes back to the slot it was pulled from:
// slot q of x - e_q, or (times alpha) slot INV[q] of x.  Every slot of lam_out has exactly
// one writer: the pull is a one-to-one map of slots; slots that only halo cells read are
// set to zero by their own cell (the halo holds the rest state, a constant).
template <int MODE>
__global__ __launch_bounds__(256) void k_adjoint(AdjArgs a) {
    using C = typename Tr<MODE>::C;
    const GridDesc &g = a.g;
    // 256-cell row segments, 32-bit tile arithmetic (as the forward sweep)
    const int tiles_x = (g.nx + 255) >> 8;
    const int b = blockIdx.x;
    const int row = b / tiles_x;
    const int i = (b - row * tiles_x) * 256 + (int)threadIdx.x + 1;
    if (i > g.nx) return;
    const int j = row % g.ny + 1;
    const int k = row / g.ny + 1;
    const long long c = cell_index(g, i, j, k);
    const long long N = g.ncell;
    C *lo = static_cast<C *>(a.lam_out);
    const uint8_t m = a.mask[c];
    if (m & MK_SOLID) {
#pragma unroll
        for (int q = 0; q < Q; ++q) lo[pop_index<MODE>(q, c, N)] = C(0);
        return;
    }
    const uint32_t sb = (m & MK_NEAR) ? a.wall_bits[c] : 0u;   // bit q: c - e_q is solid
    // u of the pulled populations, recorded by the replay's sweep (k_step_ab with VEL)
    const C *vel = static_cast<const C *>(a.vel);
    const C ux = vel[c], uy = vel[N + c], uz = vel[2 * N + c];
    const C u2 = ux * ux + uy * uy + uz * uz;
    const C sig = fmax(static_cast<const C *>(a.sig_x)[i],
                       fmax(static_cast<const C *>(a.sig_y)[j], static_cast<const C *>(a.sig_z)[k]));
    const C *li = static_cast<const C *>(a.lam_in);
    C lam[Q];
    C A = C(0), Bx = C(0), By = C(0), Bz = C(0), Cx = C(0), Cy = C(0), Cz = C(0), D = C(0);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        lam[q] = li[pop_index<MODE>(q, c, N)];
        const C wl = wq<C>(q) * lam[q];
        const C eu = (C)EX[q] * ux + (C)EY[q] * uy + (C)EZ[q] * uz;
        A += wl;
        Bx += wl * (C)EX[q]; By += wl * (C)EY[q]; Bz += wl * (C)EZ[q];
        Cx += wl * eu * (C)EX[q]; Cy += wl * eu * (C)EY[q]; Cz += wl * eu * (C)EZ[q];
        D += wl * eu * eu;
    }
    const C om = (C)a.omega, keep = C(1) - sig;
    const C base = A * (C(1) + C(1.5) * u2) - C(4.5) * D;
    const C vx = C(3) * Bx + C(9) * Cx - C(3) * A * ux;
    const C vy = C(3) * By + C(9) * Cy - C(3) * A * uy;
    const C vz = C(3) * Bz + C(9) * Cz - C(3) * A * uz;
    C probe = C(0);
    if ((m & MK_PROBE) && a.dl_dp) {
        const int u = probe_slot(a.probe_cells, a.n_probe_cells, c);
        if (u >= 0) probe = (C)a.dl_dp[u];
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const C mu = keep * ((C(1) - om) * lam[q] + om * (base + (C)EX[q] * vx + (C)EY[q] * vy + (C)EZ[q] * vz)) + probe;
        if ((sb >> q) & 1u) {
            lo[pop_index<MODE>(INV[q], c, N)] = (C)a.alpha * mu;
        } else {
            const int si = i - EX[q], sj = j - EY[q], sk = k - EZ[q];
            const int out = (si < 1 || si > g.nx) + (sj < 1 || sj > g.ny) + (sk < 1 || sk > g.nz);
            if (out == 0) {
                lo[pop_index<MODE>(q, c - nb_off(g, q), N)] = mu;
            } else if (out == 1 && a.nscbc_faces) {
                // a ghost on a face cell of an NSCBC face: its adjoint waits in the halo slot
                const int face = si < 1 ? 0 : si > g.nx ? 1 : sj < 1 ? 2 : sj > g.ny ? 3 : sk < 1 ? 4 : 5;
                if (((a.nscbc_faces >> face) & 1) && !(a.mask[c - nb_off(g, q)] & MK_SOLID))
                    lo[pop_index<MODE>(q, c - nb_off(g, q), N)] = mu;
            }
        }
        // slot q of c is read only by the halo cell c + e_q (a constant): zero it
        const int ti = i + EX[q], tj = j + EY[q], tk = k + EZ[q];
        if (q != 0 && (ti < 1 || ti > g.nx || tj < 1 || tj > g.ny || tk < 1 || tk > g.nz) &&
            !(a.mask[c + nb_off(g, q)] & MK_SOLID))
            lo[pop_index<MODE>(q, c, N)] = C(0);
    }
}

// Transposed NSCBC ghost fill of one face (05:684-835).  Forward, per face cell h of the
// halo and its interior neighbour I: (rho_I, u_I) = moments of f_I; on a minus face
// W = u_I,a - cs (rho_I - 1), rho_B = 1 - W / (2 cs), u_B,a = W / 2 (plus face: W = u_I,a +
// cs (rho_I - 1), rho_B = 1 + W / (2 cs)), tangential u_B = u_I; ghost_q = feq_q(rho_B, u_B)
// + f_I,q - feq_q(rho_I, u_I) for the slots interior cells pull.  The transpose sends the
// adjoint g_q of those slots to I:
//   lam_I,p += g_p + d_rho + e_p . d_j,
// with (d_rho, d_j) the adjoint of the moments, through
//   dF/drho = w (1 + 3 e.u + 4.5 (e.u)^2 - 1.5 u^2), dF/du = w rho (3 e + 9 (e.u) e - 3 u),
//   d rho_B/d rho_I = 1/2, d rho_B/d u_I,a = -+ 1/(2 cs), d u_B,a/d rho_I = -+ cs/2,
//   d u_B,a/d u_I,a = 1/2, u_I = j / rho_I.
// One thread per boundary cell I, which adds the transposes of every NSCBC ghost it feeds
// (up to three at an edge or corner), so no two threads write one cell.
template <class C>
__device__ __forceinline__ void feq_partials(int q, C r, C ux, C uy, C uz, C &dr, C &du0, C &du1, C &du2) {
    const C eu = (C)EX[q] * ux + (C)EY[q] * uy + (C)EZ[q] * uz;
    const C w = wq<C>(q);
    dr = w * (C(1) + C(3) * eu + C(4.5) * eu * eu - C(1.5) * (ux * ux + uy * uy + uz * uz));
    const C s = w * r;
    du0 = s * (C(3) * (C)EX[q] + C(9) * eu * (C)EX[q] - C(3) * ux);
    du1 = s * (C(3) * (C)EY[q] + C(9) * eu * (C)EY[q] - C(3) * uy);
    du2 = s * (C(3) * (C)EZ[q] + C(9) * eu * (C)EZ[q] - C(3) * uz);
}

template <int MODE>
__global__ __launch_bounds__(256) void k_adjoint_ghost(AdjArgs a) {
    using C = typename Tr<MODE>::C;
    const GridDesc &g = a.g;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    // one thread per cell of the interior's outer shell, each counted once: the x-face
    // layers, then the y-face rows between them, then the z-face cells inside both
    const long long nxs = nx > 1 ? 2 : 1, nys = ny > 1 ? 2 : 1, nzs = nz > 1 ? 2 : 1;
    const long long ix = (long long)nx > 2 ? nx - 2 : 0, iy = (long long)ny > 2 ? ny - 2 : 0;
    const long long cx = nxs * ny * nz, cy = nys * ix * nz, cz = nzs * ix * iy;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (t < cx) {
        i = (t % nxs) ? nx : 1;
        t /= nxs;
        j = (int)(t % ny) + 1;
        k = (int)(t / ny) + 1;
    } else if ((t -= cx) < cy) {
        j = (t % nys) ? ny : 1;
        t /= nys;
        i = (int)(t % ix) + 2;
        k = (int)(t / ix) + 1;
    } else if ((t -= cy) < cz) {
        k = (t % nzs) ? nz : 1;
        t /= nzs;
        i = (int)(t % ix) + 2;
        j = (int)(t / ix) + 2;
    } else {
        return;
    }
    const long long N = g.ncell;
    const long long I = cell_index(g, i, j, k);
    if (a.mask[I] & MK_SOLID) return;
    C *lam = static_cast<C *>(a.lam_out);
    // the NSCBC faces this cell borders whose halo cell is fluid
    const int at[6] = {i == 1, i == nx, j == 1, j == ny, k == 1, k == nz};
    int faces = 0;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        if (!at[f] || !((a.nscbc_faces >> f) & 1)) continue;
        const int ax = f >> 1, d = (f & 1) ? 1 : -1;
        const long long h = I + (ax == 0 ? d : ax == 1 ? (long long)d * g.px : (long long)d * g.plane);
        if (!(a.mask[h] & MK_SOLID)) faces |= 1 << f;
    }
    if (!faces) return;
    // forward boundary state from S'_n(I)
    C f[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        C v0 = ld<MODE>(a.state, pop_index<MODE>(q, I, N), a.inv_scale);
        if constexpr (MODE != M_F64) v0 += wq<C>(q);
        f[q] = v0;
    }
    C r = C(0), jx = C(0), jy = C(0), jz = C(0);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        r += f[q];
        jx += f[q] * (C)EX[q];
        jy += f[q] * (C)EY[q];
        jz += f[q] * (C)EZ[q];
    }
    const C uI[3] = {jx / r, jy / r, jz / r};
    const C cs = (C)CS;
    C acc[Q];
#pragma unroll
    for (int p = 0; p < Q; ++p) acc[p] = C(0);
    C dM0 = C(0), dJ[3] = {C(0), C(0), C(0)};
#pragma unroll
    for (int fc = 0; fc < 6; ++fc) {
        if (!((faces >> fc) & 1)) continue;
        const int axis = fc >> 1;
        const bool plus = fc & 1;
        const int hi = i + (axis == 0 ? (plus ? 1 : -1) : 0);
        const int hj = j + (axis == 1 ? (plus ? 1 : -1) : 0);
        const int hk = k + (axis == 2 ? (plus ? 1 : -1) : 0);
        const long long h = cell_index(g, hi, hj, hk);
        // adjoint of the pulled ghost slots (target inside the interior), cleared once read
        C gl[Q];
        bool any = false;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            gl[q] = C(0);
            const int ti = hi + EX[q], tj = hj + EY[q], tk = hk + EZ[q];
            if (q == 0 || ti < 1 || ti > nx || tj < 1 || tj > ny || tk < 1 || tk > nz) continue;
            const long long idx = pop_index<MODE>(q, h, N);
            gl[q] = lam[idx];
            lam[idx] = C(0);
            any = any || gl[q] != C(0);
        }
        if (!any) continue;
        const C sg = plus ? C(1) : C(-1);
        const C W = uI[axis] + sg * cs * (r - C(1));
        const C rB = C(1) + sg * W / (C(2) * cs);
        C uB[3] = {uI[0], uI[1], uI[2]};
        uB[axis] = W / C(2);
        C aB = 0, vB[3] = {0, 0, 0}, aI = 0, vI[3] = {0, 0, 0};
#pragma unroll
        for (int q = 1; q < Q; ++q) {
            if (gl[q] == C(0)) continue;
            C dr, d0, d1, d2;
            feq_partials<C>(q, rB, uB[0], uB[1], uB[2], dr, d0, d1, d2);
            aB += gl[q] * dr; vB[0] += gl[q] * d0; vB[1] += gl[q] * d1; vB[2] += gl[q] * d2;
            feq_partials<C>(q, r, uI[0], uI[1], uI[2], dr, d0, d1, d2);
            aI += gl[q] * dr; vI[0] += gl[q] * d0; vI[1] += gl[q] * d1; vI[2] += gl[q] * d2;
        }
        // through rho_B(rho_I, u_I,a), u_B,a(rho_I, u_I,a) and u_B,t = u_I,t
        const C dRho = aB * C(0.5) + vB[axis] * (sg * cs * C(0.5)) - aI;
        C dU[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) dU[c] = vB[c] - vI[c];
        dU[axis] = aB * (sg / (C(2) * cs)) + vB[axis] * C(0.5) - vI[axis];
        // through u_I = j / rho_I
        const C ir = C(1) / r;
        dM0 += dRho - (dU[0] * uI[0] + dU[1] * uI[1] + dU[2] * uI[2]) * ir;
#pragma unroll
        for (int c = 0; c < 3; ++c) dJ[c] += dU[c] * ir;
#pragma unroll
        for (int p = 0; p < Q; ++p) acc[p] += gl[p];
    }
#pragma unroll
    for (int p = 0; p < Q; ++p) {
        const long long idx = pop_index<MODE>(p, I, N);
        lam[idx] += acc[p] + dM0 + (C)EX[p] * dJ[0] + (C)EY[p] * dJ[1] + (C)EZ[p] * dJ[2];
    }
}

// out[e] = sum_q w_q lam[q][cells[e]] for the n source entries of one step.
template <class T>
__global__ __launch_bounds__(256) void k_zero(T *p, long long n) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = T(0);
}

template <int MODE>
__global__ __launch_bounds__(256) void k_source_dot(const void *lam, long long ncell, const long long *cells, int n,
                                                    double *out) {
    using C = typename Tr<MODE>::C;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const C *l = static_cast<const C *>(lam);
    double acc = 0.0;
    for (int q = 0; q < Q; ++q) acc += W64[q] * (double)l[pop_index<MODE>(q, cells[e], ncell)];
    out[e] = acc;
}

// ---- reading f* of an interior cell in the current layout ---------------------------------

template <int MODE, int LAY>
__device__ __forceinline__ typename Tr<MODE>::C read_post(const void *pop, const uint8_t *mask,
                                                          const GridDesc &g, long long c, int q,
                                                          float inv_scale) {
    const long long N = g.ncell;
    if constexpr (MODE == M_B16) {
        static_assert(LAY == LAY_AB, "B16 runs with the two-array pattern only");
        return b16_slot_g(pop, N, c, q);
    } else if constexpr (LAY == LAY_AB) {
        return ld<MODE>(pop, pop_index<MODE>(q, c, N), inv_scale);
    } else if constexpr (LAY == LAY_P) {
        return ld<MODE>(pop, pop_index<MODE>(INV[q], c, N), inv_scale);
    } else {
        const long long t = c + nb_off(g, q);
        const long long idx = (mask[t] & MK_SOLID) ? pop_index<MODE>(INV[q], c, N) : pop_index<MODE>(q, t, N);
        return ld<MODE>(pop, idx, inv_scale);
    }
}

// ---- ghost layer (05:671-904 at one rank) --------------------------------------------------

template <int MODE, int LAY>
__device__ __forceinline__ void ghost_body(const GhostArgs &a, long long t) {
    using C = typename Tr<MODE>::C;
    if (t >= a.n_halo) return;
    const GridDesc &g = a.g;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    int i, j, k;
    // Halo enumeration: the z layers (k = 0 unless it is solid ground, then k = nz + 1) and the
    // two y rows of every interior k, their cells i in [1, nx] by rows of W = roundup(nx, 64)
    // lanes, so that every wave covers one 64-cell-aligned run (interior cells start
    // 64-aligned at i = 1) and its loads and stores fill whole lines; then the same rows'
    // cells i = 0 and i = nx + 1; then the two x columns of every interior (j, k).
    const long long W = ghost_row_lanes(nx);
    const int k0 = a.skip_zm ? 1 : 0;
    const long long zrows = (long long)(2 - k0) * (ny + 2);
    const long long rows = zrows + 2LL * nz;
    long long row;
    if (t < rows * W) {
        row = t / W;
        i = (int)(t - row * W) + 1;
        if (i > nx) return;
    } else if ((t -= rows * W) < 2 * rows) {
        row = t >> 1;
        i = (t & 1) ? nx + 1 : 0;
    } else {
        t -= 2 * rows;
        const long long xcol = (long long)ny * nz;
        i = t < xcol ? 0 : nx + 1;
        const long long r = t < xcol ? t : t - xcol;
        k = (int)(r / ny) + 1;
        j = (int)(r % ny) + 1;
        row = -1;
    }
    if (row >= 0) {
        if (row < zrows) {
            k = ((int)(row / (ny + 2)) + k0) == 0 ? 0 : nz + 1;
            j = (int)(row % (ny + 2));
        } else {
            const long long r = row - zrows;
            j = (r & 1) ? ny + 1 : 0;
            k = (int)(r >> 1) + 1;
        }
    }
    const long long N = g.ncell;
    const long long h = cell_index(g, i, j, k);
    // The halo cell's mask, the interior neighbour's mask and the interior neighbour's
    // populations are loaded together (their addresses need no memory); a solid halo cell
    // returns, a solid neighbour or (LAY_Q) a neighbour next to a wall then overrides the
    // populations.  One memory round trip instead of three: at 1024x1024x256 the kernel
    // takes 43.9 instead of 45.3 us (FP16S) and 93.6 instead of 95.7 us (FP64).
    const uint8_t mh = a.mask[h];

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
    const int ii = i - oi, jj = j - oj, kk = k - ok;
    const long long I = cell_index(g, ii, jj, kk);
    uint8_t mI = 0;
    if (ns) {
        mI = a.mask[I];
        // the populations of a fluid neighbour, in every layout
        // constant indices only: a scene's arguments copied into registers (k_ghost_multi)
        // would otherwise go to private memory for a runtime index
        const void *fbf = (face & 1) ? a.fb[1] : a.fb[0];
        if (axis == 0 && fbf) {
            const long long r0 = ((long long)(jj - 1) + (long long)ny * (kk - 1)) * Q;
            if constexpr (MODE == M_B16) {
                const float *fb = static_cast<const float *>(fbf) + r0;
#pragma unroll
                for (int q = 0; q < Q; ++q) fI[q] = fb[q];
            } else {
                const typename Tr<MODE>::S *fb = static_cast<const typename Tr<MODE>::S *>(fbf) + r0;
#pragma unroll
                for (int q = 0; q < Q; ++q) fI[q] = ld<MODE>(fb, q, a.inv_scale);   // per-thread address
            }
        } else if constexpr (MODE == M_B16) {
            b16_load_g(a.pop, g.ncell, I, fI);
        } else {
            // f*_q(I) in slot q of I (LAY_AB), slot INV[q] of I (LAY_P), or, with no solid
            // neighbour, slot q of I + e_q (LAY_Q)
#pragma unroll
            for (int q = 0; q < Q; ++q)
                fI[q] = ld<MODE>(a.pop, LAY == LAY_AB ? pop_index<MODE>(q, I, N)
                                        : LAY == LAY_P ? pop_index<MODE>(INV[q], I, N)
                                                       : pop_index<MODE>(q, I + nb_off(g, q), N), a.inv_scale);
        }
    }
    if (mh & MK_SOLID) return;
    if (ns) {
        if (mI & MK_SOLID) {
            int pidx;
            if (axis == 0) pidx = (jj - 1) + ny * (kk - 1);
            else if (axis == 1) pidx = (ii - 1) + nx * (kk - 1);
            else pidx = (ii - 1) + nx * (jj - 1);
            const int *fm = face == 0 ? a.face_map[0] : face == 1 ? a.face_map[1] : face == 2 ? a.face_map[2]
                          : face == 3 ? a.face_map[3] : face == 4 ? a.face_map[4] : a.face_map[5];
            const int slot = fm[pidx];
            const C *so = static_cast<const C *>(a.s_old);
#pragma unroll
            for (int q = 0; q < Q; ++q) fI[q] = so[(long long)slot * Q + q];
        } else if constexpr (MODE != M_B16 && LAY == LAY_Q) {
            if (mI & MK_NEAR) {
#pragma unroll
                for (int q = 0; q < Q; ++q) fI[q] = read_post<MODE, LAY>(a.pop, a.mask, g, I, q, a.inv_scale);
            }
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

    C gout[Q];            // B16: the cell is coded once, after the loop
    unsigned pulled = 0;  // B16: slots that interior cells pull
    // x-halo cell whose pulled slots go to the records (see StepArgs::xh)
    const bool xcell = LAY == LAY_AB && a.xh && (i == 0 || i == nx + 1);
#pragma unroll
    for (int q = 0; q < Q; ++q) gout[q] = rest<MODE>(q);
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
        if constexpr (MODE == M_B16) {
            gout[q] = val;
            pulled |= 1u << q;
            continue;
        } else {
            if (xcell) {
                st<MODE>(a.xh, xh_index(g, q, j, k), val, a.scale);
                continue;
            }
            const long long tc = h + nb_off(g, q);
            long long idx;
            if constexpr (LAY == LAY_AB) {
                idx = pop_index<MODE>(q, h, N);
            } else if constexpr (LAY == LAY_P) {
                idx = pop_index<MODE>(INV[q], h, N);
            } else {
                idx = (a.mask[tc] & MK_SOLID) ? pop_index<MODE>(INV[q], h, N) : pop_index<MODE>(q, tc, N);
            }
            st<MODE>(a.pop, idx, val, a.scale);
        }
    }
    if constexpr (MODE == M_B16) {
        // slots no interior cell pulls are never read: only the pulled ones are written
        float x[Q];
        b16_g_to_x(gout, x);
        if (xcell) {
            // the coding of b16_store_x, into the records
            const int e = b16_exp_of(x);
            short m[Q];
            b16_code(x, e, m);
#pragma unroll
            for (int q = 1; q < Q; ++q)
                if ((pulled >> q) & 1u) static_cast<short *>(a.xh)[xh_index(g, q, j, k)] = m[q];
            static_cast<signed char *>(a.xh)[xh_exp_offset(g, i == 0 ? 0 : 1, j, k)] = (signed char)(e - B16_BITS);
        } else {
            b16_store_x(a.pop, N, h, x, pulled);
        }
    }
}

template <int MODE, int LAY>
__global__ __launch_bounds__(256) void k_ghost(GhostArgs a) {
    // the step counter moves on here, after the step's source adds read it and before its
    // sweep does (the sweep subtracts StepArgs::ctr_lag): one launch fewer per step
    if (a.bump && blockIdx.x == 0 && threadIdx.x == 0) *a.bump += 1;
    ghost_body<MODE, LAY>(a, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// Several scenes in one launch: workgroup g serves scene blocks[g].x, items
// blocks[g].y .. blocks[g].y + 255 of it (each scene's range starts on a new workgroup, so
// the scene is uniform over the workgroup and its arguments are scalar loads).
template <int MODE, int LAY>
__global__ __launch_bounds__(256) void k_ghost_multi(const GhostArgs *__restrict__ scenes,
                                                     const int2 *__restrict__ blocks) {
    const int2 sb = blocks[blockIdx.x];
    const GhostArgs a = scenes[sb.x];
    ghost_body<MODE, LAY>(a, (long long)sb.y + threadIdx.x);
}

// ---- solid cells on open faces (05:917-938 for a SOLID dst) -------------------------------
// f_next[b][q] = solid(b - e_q) ? f[b][inv q] : f*_q(b - e_q).  With b solid, f*_q of a
// fluid neighbour sits in slot INV[q] of that neighbour in either A-A layout and in slot q
// of it in LAY_AB; the ghost kernel puts a halo neighbour's value in the same place.
template <int MODE>
__device__ __forceinline__ void solid_body(const SolidArgs &a, int n) {
    using C = typename Tr<MODE>::C;
    if (n >= a.n) return;
    const long long c = a.cells[n];
    const long long N = a.g.ncell;
    const C *so = static_cast<const C *>(a.s_old);
    C *sn = static_cast<C *>(a.s_new);
    // the cell's (i, j, k), once: its source cell of slot q is (i, j, k) - e_q
    const long long r = c - a.g.origin;
    const int ci = (int)(r % a.g.px), cj = (int)((r / a.g.px) % a.g.py), ck = (int)(r / a.g.plane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const long long src = c - nb_off(a.g, q);
        C v;
        // an x-halo source keeps this slot in the records (see StepArgs::xh)
        const bool xsrc = a.xh && EX[q] != 0 && (EX[q] > 0 ? ci - 1 == 0 : ci + 1 == a.g.nx + 1);
        const int sj = cj - EY[q], sk = ck - EZ[q];
        if (a.mask[src] & MK_SOLID) {
            v = so[(long long)n * Q + INV[q]];
            if (a.alpha != 1.0) v = reflect<MODE>(q, v, a.alpha);
        } else if (xsrc) {
            if constexpr (MODE == M_B16) {
                const int E = static_cast<const signed char *>(a.xh)[xh_exp_offset(a.g, EX[q] > 0 ? 0 : 1, sj, sk)];
                v = b16_dec(static_cast<const short *>(a.xh)[xh_index(a.g, q, sj, sk)], E) * wq<float>(q);
            } else {
                v = ld<MODE>(a.xh, xh_index(a.g, q, sj, sk), a.inv_scale);
            }
        } else if constexpr (MODE == M_B16) {
            v = b16_slot_g(a.pop, N, src, q);
        } else {
            v = ld<MODE>(a.pop, pop_index<MODE>(a.lay == LAY_AB ? q : INV[q], src, N), a.inv_scale);
        }
        sn[(long long)n * Q + q] = v;
    }
}

template <int MODE>
__global__ __launch_bounds__(256) void k_solid(SolidArgs a) {
    solid_body<MODE>(a, blockIdx.x * blockDim.x + threadIdx.x);
}

template <int MODE>
__global__ __launch_bounds__(256) void k_solid_multi(const SolidArgs *__restrict__ scenes,
                                                     const int2 *__restrict__ blocks) {
    const int2 sb = blocks[blockIdx.x];
    const SolidArgs a = scenes[sb.x];
    solid_body<MODE>(a, sb.y + (int)threadIdx.x);
}

// ---- source injection (05:660-665) ---------------------------------------------------------
template <int MODE, int LAY>
__device__ __forceinline__ void source_body(const SourceArgs &a, int e, int q) {
    const long long N = a.g.ncell;
    const long long c = a.cells[e];
    long long idx;
    if constexpr (LAY == LAY_AB) {
        idx = pop_index<MODE>(q, c, N);
    } else if constexpr (LAY == LAY_P) {
        idx = pop_index<MODE>(INV[q], c, N);
    } else {
        const long long tc = c + nb_off(a.g, q);
        idx = (a.mask[tc] & MK_SOLID) ? pop_index<MODE>(INV[q], c, N) : pop_index<MODE>(q, tc, N);
    }
    typename Tr<MODE>::C v;
    if constexpr (MODE == M_F64) {
        double *p = static_cast<double *>(a.pop);
        p[idx] += W64[q] * a.adds[e];
        v = p[idx];
    } else {
        v = ld<MODE>(a.pop, idx, a.inv_scale) + (float)(W64[q] * a.adds[e]);
        st<MODE>(a.pop, idx, v, a.scale);
    }
    if constexpr (LAY == LAY_P) {
        // keep the wall copy of a population headed into a wall in step (see Layout)
        const long long tc = c + nb_off(a.g, q);
        if (q != 0 && (a.mask[tc] & MK_SOLID))
            st<MODE>(a.pop, pop_index<MODE>(q, tc, N), a.alpha != 1.0 ? reflect<MODE>(q, v, a.alpha) : v, a.scale);
    }
    // the x-face buffer holds the same f*_q of boundary cells; apply the same add to it
    const long long r = c - a.g.origin;
    const int i = (int)(r % a.g.px);
    if ((i == 1 && a.fb[0]) || (i == a.g.nx && a.fb[1])) {
        const int j = (int)((r / a.g.px) % a.g.py);
        const int k = (int)(r / a.g.plane);
        const long long fidx = ((long long)(j - 1) + (long long)a.g.ny * (k - 1)) * Q + q;
        for (int face = 0; face < 2; ++face) {
            if ((face == 0 ? i != 1 : i != a.g.nx) || !a.fb[face]) continue;
            if constexpr (MODE == M_F64) {
                static_cast<double *>(a.fb[face])[fidx] += W64[q] * a.adds[e];
            } else {
                const float v = ld<MODE>(a.fb[face], fidx, a.inv_scale) + (float)(W64[q] * a.adds[e]);
                st<MODE>(a.fb[face], fidx, v, a.scale);
            }
        }
    }
}

// B16 (two-array pattern): one thread per stencil cell adds to all 19 slots and re-codes
// the cell; g_q += w_q * add is x_q += add in normalised units.
__device__ __forceinline__ void source_cell_b16(const SourceArgs &a, int e) {
    const long long c = a.cells[e];
    const long long N = a.g.ncell;
    const float add = (float)a.adds[e];
    float x[Q];
    b16_load_g(a.pop, N, c, x);
    b16_g_to_x(x, x);
#pragma unroll
    for (int q = 0; q < Q; ++q) x[q] += add;
    const int ex = b16_store_x(a.pop, N, c, x);
    // an x-face record of this cell takes the stored values (as the sweep writes them)
    const long long r = c - a.g.origin;
    const int i = (int)(r % a.g.px);
    for (int face = 0; face < 2; ++face) {
        if ((face == 0 ? i != 1 : i != a.g.nx) || !a.fb[face]) continue;
        const int j = (int)((r / a.g.px) % a.g.py);
        const int k = (int)(r / a.g.plane);
        float *rec = static_cast<float *>(a.fb[face]) + ((long long)(j - 1) + (long long)a.g.ny * (k - 1)) * Q;
        short m[Q];
        b16_code(x, ex, m);
        rec[0] = x[0] * wq<float>(0);
#pragma unroll
        for (int q = 1; q < Q; ++q) rec[q] = b16_dec(m[q], ex - B16_BITS) * wq<float>(q);
    }
}

template <int MODE, int LAY>
__global__ __launch_bounds__(256) void k_source(SourceArgs a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if constexpr (MODE == M_B16) {
        if (t < a.n) source_cell_b16(a, t);
    } else {
        if (t >= a.n * Q) return;
        source_body<MODE, LAY>(a, t / Q, t % Q);
    }
}

// All scenes' source entries of the current step (index ctr - ctr_base into step_off).
template <int MODE, int LAY>
__global__ __launch_bounds__(256) void k_source_multi(SourceMulti m) {
    const long long st = *m.ctr - *m.ctr_base;
    const long long e0 = m.step_off[st], e1 = m.step_off[st + 1];
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    constexpr int PER = MODE == M_B16 ? 1 : Q;   // threads per entry
    if (t >= (e1 - e0) * PER) return;
    const long long e = e0 + t / PER;
    SourceArgs a = m.scenes[m.entry_scene ? m.entry_scene[e] : 0];   // null: one scene
    a.cells = m.cells + e;
    a.adds = m.adds + e;
    a.n = 1;
    if constexpr (MODE == M_B16) source_cell_b16(a, 0);
    else source_body<MODE, LAY>(a, 0, (int)(t % Q));
}

// Fills the x-face buffers from the lattice (after the state was set from outside).
template <int MODE, int LAY>
__global__ __launch_bounds__(256) void k_fb_gather(const void *pop, const uint8_t *mask, GridDesc g,
                                                  void *fb0, void *fb1, float scale, float inv_scale) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nface = (long long)g.ny * g.nz;
    if (t >= 2 * nface) return;
    const int face = t < nface ? 0 : 1;
    void *fb = face == 0 ? fb0 : fb1;
    if (!fb) return;
    const long long jk = face == 0 ? t : t - nface;
    const int j = (int)(jk % g.ny) + 1, k = (int)(jk / g.ny) + 1;
    const long long I = cell_index(g, face == 0 ? 1 : g.nx, j, k);
    if (mask[I] & MK_SOLID) return;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const auto v = read_post<MODE, LAY>(pop, mask, g, I, q, inv_scale);
        if constexpr (MODE == M_B16) static_cast<float *>(fb)[jk * Q + q] = v;
        else st<MODE>(fb, jk * Q + q, v, scale);
    }
}

// ---- initialisation and gathers ------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(256) void k_fill_rest(void *pop, long long ncell) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell) return;
    if constexpr (MODE == M_B16) {
        b16_rest(pop)[t] = 0.0f;
#pragma unroll
        for (int q = 1; q < Q; ++q) b16_mants(pop, ncell)[mant_index(q, t)] = 0;
        b16_exps(pop, ncell)[t] = 0;
        return;
    } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) st<MODE>(pop, pop_index<MODE>(q, t, ncell), rest<MODE>(q), 1.0f);
    }
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
    const long long c = cell_index(g, i, j, k);
    if (mask[c] & MK_SOLID) return;
    const double r = rho[n];
    if constexpr (MODE == M_B16) {
        float x[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) x[q] = (float)(r - 1.0);   // g = w (rho - 1)
        b16_store_x(pop, g.ncell, c, x);
        return;
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        // feq(q, rho, 0, 0, 0) = w_q * rho (05:286-291 with u = 0); symmetric in q, so the
        // value is the same in either slot convention.
        if constexpr (MODE == M_F64) st<MODE>(pop, pop_index<MODE>(q, c, g.ncell), W64[q] * r, scale);
        else if constexpr (MODE != M_B16) st<MODE>(pop, pop_index<MODE>(q, c, g.ncell), (float)(W64[q] * (r - 1.0)), scale);
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
    const long long c = cell_index(g, i, j, k);
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
    if constexpr (MODE == M_B16) {
        return hipErrorNotSupported;   // two-array pattern only
    } else {
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
}

template <int MODE>
hipError_t Launch<MODE>::step_ab_vel(const StepArgs &a, dim3 block, hipStream_t s) {
    if constexpr (MODE == M_B16 || MODE == M_F16) {
        return hipErrorNotSupported;
    } else {
        k_step_ab<MODE, false, true><<<dim3(a.nblocks), block, 0, s>>>(a);
        return hipGetLastError();
    }
}

template <int MODE>
hipError_t Launch<MODE>::step_ab(const StepArgs &a, bool capture, dim3 block, hipStream_t s) {
    const dim3 grid(a.nblocks);
    if (capture) k_step_ab<MODE, true><<<grid, block, 0, s>>>(a);
    else k_step_ab<MODE, false><<<grid, block, 0, s>>>(a);
    return hipGetLastError();
}

template <int MODE>
hipError_t Launch<MODE>::ghost(const GhostArgs &a, int lay, hipStream_t s) {
    if (a.n_halo <= 0) return hipSuccess;
    const unsigned nb = blocks_for(a.n_halo, 256);
    if constexpr (MODE == M_B16) {
        if (lay != LAY_AB) return hipErrorNotSupported;
        k_ghost<MODE, LAY_AB><<<nb, 256, 0, s>>>(a);
    } else {
        if (lay == LAY_P) k_ghost<MODE, LAY_P><<<nb, 256, 0, s>>>(a);
        else if (lay == LAY_Q) k_ghost<MODE, LAY_Q><<<nb, 256, 0, s>>>(a);
        else k_ghost<MODE, LAY_AB><<<nb, 256, 0, s>>>(a);
    }
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
    const unsigned nb = blocks_for((long long)a.n * (MODE == M_B16 ? 1 : Q), 256);
    if constexpr (MODE == M_B16) {
        if (lay != LAY_AB) return hipErrorNotSupported;
        k_source<MODE, LAY_AB><<<nb, 256, 0, s>>>(a);
    } else {
        if (lay == LAY_P) k_source<MODE, LAY_P><<<nb, 256, 0, s>>>(a);
        else if (lay == LAY_Q) k_source<MODE, LAY_Q><<<nb, 256, 0, s>>>(a);
        else k_source<MODE, LAY_AB><<<nb, 256, 0, s>>>(a);
    }
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
hipError_t Launch<MODE>::prime_walls(void *pop, const uint8_t *mask, const GridDesc &g, int lay, double alpha,
                                     hipStream_t s) {
    const long long total = (long long)g.nx * g.ny * g.nz;
    if (lay != LAY_P) return hipErrorInvalidValue;
    if constexpr (MODE == M_B16) {
        return hipErrorNotSupported;
    } else {
        k_prime_walls<MODE><<<blocks_for(total, 256), 256, 0, s>>>(pop, mask, g, alpha);
        return hipGetLastError();
    }
}

template <int MODE>
hipError_t Launch<MODE>::fb_gather(const void *pop, const uint8_t *mask, const GridDesc &g, int lay,
                                   void *fb0, void *fb1, float scale, float inv_scale, hipStream_t s) {
    const long long n = 2LL * g.ny * g.nz;
    if constexpr (MODE == M_B16) {
        if (lay != LAY_AB) return hipErrorNotSupported;
        k_fb_gather<MODE, LAY_AB><<<blocks_for(n, 256), 256, 0, s>>>(pop, mask, g, fb0, fb1, scale, inv_scale);
    } else {
        if (lay == LAY_P) k_fb_gather<MODE, LAY_P><<<blocks_for(n, 256), 256, 0, s>>>(pop, mask, g, fb0, fb1, scale, inv_scale);
        else if (lay == LAY_Q) k_fb_gather<MODE, LAY_Q><<<blocks_for(n, 256), 256, 0, s>>>(pop, mask, g, fb0, fb1, scale, inv_scale);
        else k_fb_gather<MODE, LAY_AB><<<blocks_for(n, 256), 256, 0, s>>>(pop, mask, g, fb0, fb1, scale, inv_scale);
    }
    return hipGetLastError();
}

template <int MODE>
hipError_t Launch<MODE>::step_ab_walk(const StepArgs *scenes, const int2 *items, int n_items, int grid, hipStream_t s) {
    if (n_items <= 0) return hipSuccess;
    k_step_ab_walk<MODE><<<std::min(grid, n_items), 256, 0, s>>>(scenes, items, n_items);
    return hipGetLastError();
}

template <int MODE>
hipError_t Launch<MODE>::step_ab_multi(const StepArgs *scenes, const int *n_tiles, int n_scenes, int max_tiles,
                                       hipStream_t s) {
    if (n_scenes <= 0 || max_tiles <= 0) return hipSuccess;
    k_step_ab_multi<MODE><<<dim3(max_tiles, n_scenes), 256, 0, s>>>(scenes, n_tiles);
    return hipGetLastError();
}

template <int MODE>
hipError_t Launch<MODE>::ghost_multi(const GhostArgs *scenes, const int2 *blocks, int n_blocks, hipStream_t s) {
    if (n_blocks <= 0) return hipSuccess;
    k_ghost_multi<MODE, LAY_AB><<<n_blocks, 256, 0, s>>>(scenes, blocks);
    return hipGetLastError();
}

template <int MODE>
hipError_t Launch<MODE>::solid_multi(const SolidArgs *scenes, const int2 *blocks, int n_blocks, hipStream_t s) {
    if (n_blocks <= 0) return hipSuccess;
    k_solid_multi<MODE><<<n_blocks, 256, 0, s>>>(scenes, blocks);
    return hipGetLastError();
}

template <int MODE>
hipError_t Launch<MODE>::source_multi(const SourceMulti &m, long long max_entries, hipStream_t s) {
    if (max_entries <= 0) return hipSuccess;
    k_source_multi<MODE, LAY_AB><<<blocks_for(max_entries * (MODE == M_B16 ? 1 : Q), 256), 256, 0, s>>>(m);
    return hipGetLastError();
}

template <int MODE>
hipError_t Launch<MODE>::adjoint(const AdjArgs &a, hipStream_t s) {
    if constexpr (MODE == M_B16 || MODE == M_F16) {
        return hipErrorNotSupported;
    } else {
        const long long tiles = (long long)((a.g.nx + 255) >> 8) * a.g.ny * a.g.nz;
        k_adjoint<MODE><<<(unsigned)tiles, 256, 0, s>>>(a);
        return hipGetLastError();
    }
}

template <int MODE>
hipError_t Launch<MODE>::adjoint_ghost(const AdjArgs &a, hipStream_t s) {
    if constexpr (MODE == M_B16 || MODE == M_F16) {
        return hipErrorNotSupported;
    } else {
        const GridDesc &g = a.g;
        const long long nxs = g.nx > 1 ? 2 : 1, nys = g.ny > 1 ? 2 : 1, nzs = g.nz > 1 ? 2 : 1;
        const long long ix = g.nx > 2 ? g.nx - 2 : 0, iy = g.ny > 2 ? g.ny - 2 : 0;
        const long long n = nxs * g.ny * g.nz + nys * ix * g.nz + nzs * ix * iy;
        k_adjoint_ghost<MODE><<<blocks_for(n, 256), 256, 0, s>>>(a);
        return hipGetLastError();
    }
}

template <int MODE>
hipError_t Launch<MODE>::source_dot(const void *lam, long long ncell, const long long *cells, int n, double *out,
                                    hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_source_dot<MODE><<<blocks_for(n, 256), 256, 0, s>>>(lam, ncell, cells, n, out);
    return hipGetLastError();
}

template <int MODE>
hipError_t Launch<MODE>::fill_compute(void *p, long long n, hipStream_t s) {
    using C = typename Tr<MODE>::C;
    if (n <= 0) return hipSuccess;
    k_zero<C><<<blocks_for(n, 256), 256, 0, s>>>(static_cast<C *>(p), n);
    return hipGetLastError();
}

template <int MODE>
hipError_t Launch<MODE>::gather_f(const void *pop, const uint8_t *mask, const GridDesc &g, int lay,
                                  float inv_scale, double *out, hipStream_t s) {
    const long long total = (long long)g.nx * g.ny * g.nz;
    if constexpr (MODE == M_B16) {
        if (lay != LAY_AB) return hipErrorNotSupported;
        k_gather_f<MODE, LAY_AB><<<blocks_for(total, 256), 256, 0, s>>>(pop, mask, g, inv_scale, out);
    } else {
        if (lay == LAY_P) k_gather_f<MODE, LAY_P><<<blocks_for(total, 256), 256, 0, s>>>(pop, mask, g, inv_scale, out);
        else if (lay == LAY_Q) k_gather_f<MODE, LAY_Q><<<blocks_for(total, 256), 256, 0, s>>>(pop, mask, g, inv_scale, out);
        else k_gather_f<MODE, LAY_AB><<<blocks_for(total, 256), 256, 0, s>>>(pop, mask, g, inv_scale, out);
    }
    return hipGetLastError();
}

}  // namespace nlos
