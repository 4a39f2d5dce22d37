// stage_kernel.hpp -- the fused DG-SEM Maxwell RK stage for gfx950.
//
// One launch = one stage of cem_maxwell_op_rk (src/cem_maxwell.F:327-345):
// volume curl (maxwell_wght_curl :1428-1497 over local_grad3,
// src/nek5_grad.F:2-19), face restriction (:604-652), on-device face exchange
// (replaces gs_op_fields(gsh_face,...) :962), PEC mirror (:1368-1426),
// upwind/central flux (:922-1002), lift (:654-756), inverse mass
// (:1833-1904) and the 2N-storage RK update (rk4_upd,
// src/cem_common.F:18-76), fused so every field value crosses HBM once per
// stage in each direction.
//
// Mapping: one workgroup per element; lane (i,j) of the n1 x n1 lanes owns
// the column (i,j,0..n1-1) of all six fields in registers.  The element is
// also staged in LDS so lanes can read rows (r-derivative) and columns
// (s-derivative); the t-derivative runs down the lane's own registers.  For
// the four x/y faces the lanes are re-mapped onto face nodes, the fluxes go
// through LDS back to the owning lanes; the two z faces are computed by the
// owners directly.  Geometry is per element (axis-aligned affine hexahedra):
// g = (rx, sy, tz)/J and a lift factor per face (area / (J w3) at the face
// node, constant on an affine face).
//
// Field order in HBM: 0 Hx, 1 Hy, 2 Hz, 3 Ex, 4 Ey, 5 Ez, each a contiguous
// array of npts doubles at stride fs, volume index i + n1*j + n1^2*k +
// n1^3*e as in the reference.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nkc {

// Kernel face order (symmetric): 0 x-, 1 x+, 2 y-, 3 y+, 4 z-, 5 z+.
enum FaceType : int32_t { FACE_PEC = -1, FACE_NONE = -2 };

enum StageMode : int { MODE_RK = 0, MODE_RESIDUAL = 1 };
enum MatMode : int { MAT_ELEM = 0, MAT_NODE = 1 };

struct StageArgs {
    const double *__restrict__ uin;   // 6 fields, stride fs
    double *__restrict__ uout;        // MODE_RK: updated fields
    double *__restrict__ kbuf;        // MODE_RK: RK register, in place
    double *__restrict__ res;         // MODE_RESIDUAL: residual out
    const double *__restrict__ D;     // n1*n1, D(i,l) at D[i + n1*l]
    const double *__restrict__ geo;   // per element: gx gy gz L0..L5 (9)
    const int32_t *__restrict__ nbr;  // per element: 6 neighbour ids / FaceType
    const double *__restrict__ fco;   // ELEM: e*6*4 ; NODE: (e*6+f)*n2*4
    const double *__restrict__ ieps;  // ELEM: per element ; NODE: per node
    const double *__restrict__ imu;
    int64_t fs;                       // field stride (doubles)
    int32_t nelt;
    int32_t per_xcd;                  // elements per XCD slice (grid remap)
    double rka, rkb, dt;
};

// n x v for n = sg * e_AX.
template <int AX>
__device__ __forceinline__ void ncross(double sg, const double v[3], double o[3]) {
    if (AX == 0) { o[0] = 0.0;        o[1] = -sg * v[2]; o[2] = sg * v[1]; }
    if (AX == 1) { o[0] = sg * v[2];  o[1] = 0.0;        o[2] = -sg * v[0]; }
    if (AX == 2) { o[0] = -sg * v[1]; o[1] = sg * v[0];  o[2] = 0.0; }
}

// Flux at one face node (cem_maxwell_flux3d loops 1-2 + gs + PEC), scaled by
// the lift factor L.  um/up: own/neighbour fields in HBM order (H then E).
// Returns FH (lift into resH) and FE (lift into resE).
template <int AX>
__device__ __forceinline__ void face_flux(double sg, int type, const double um[6],
                                          const double up[6], const double cf[4],
                                          double L, double FH[3], double FE[3]) {
    double sE[3], sH[3];
    if (type >= 0) {
        // n x [E] = (-n x E-) + (-n' x E+), n' = -n
        double dE[3] = {up[3] - um[3], up[4] - um[4], up[5] - um[5]};
        double dH[3] = {up[0] - um[0], up[1] - um[1], up[2] - um[2]};
        ncross<AX>(sg, dE, sE);
        ncross<AX>(sg, dH, sH);
    } else {
        // no partner: gs leaves -n x E-, -n x H-; PEC doubles E, zeroes H
        double Em[3] = {-um[3], -um[4], -um[5]};
        double Hm[3] = {-um[0], -um[1], -um[2]};
        ncross<AX>(sg, Em, sE);
        if (type == FACE_PEC) {
            sE[0] *= 2.0; sE[1] *= 2.0; sE[2] *= 2.0;
            sH[0] = 0.0; sH[1] = 0.0; sH[2] = 0.0;
        } else {
            ncross<AX>(sg, Hm, sH);
        }
    }
    double fu[3], fw[3];
    ncross<AX>(sg, sH, fu);
    ncross<AX>(sg, sE, fw);
    // cf = (Y02, C02Y, Z02, C02Z)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        FH[c] = L * (cf[0] * sE[c] - cf[1] * fu[c]);
        FE[c] = L * (cf[2] * sH[c] - cf[3] * fw[c]);
    }
}

__device__ __forceinline__ int remap_element(int b, int nelt, int per_xcd) {
    // Workgroups are dealt round-robin to the 8 XCDs; give each XCD a
    // contiguous run of elements so x/y neighbours share its L2.
    int xcd = b & 7;
    int slot = b >> 3;
    return xcd * per_xcd + slot;
}

template <int N1, int MAT>
__device__ __forceinline__ void load_coef(const StageArgs &a, int e, int f, int q, double c4[4]) {
    const double *p = (MAT == MAT_ELEM) ? a.fco + ((int64_t)e * 6 + f) * 4
                                        : a.fco + (((int64_t)e * 6 + f) * (N1 * N1) + q) * 4;
    c4[0] = p[0]; c4[1] = p[1]; c4[2] = p[2]; c4[3] = p[3];
}

// LDS footprint (doubles) of one element: s_u (6 n1^3), s_f (4 faces x 6
// components x n1^2), s_D (n1^2).
template <int N1>
struct StageLds {
    static constexpr int N2 = N1 * N1, N3 = N2 * N1;
    static constexpr int U = 0, F = 6 * N3, Dm = F + 24 * N2, TOTAL = Dm + N2;
};

template <int N1, int MAT, int MODE>
__global__ void __launch_bounds__(((N1 * N1 + 63) / 64) * 64)
stage_kernel(StageArgs a) {
    constexpr int N2 = N1 * N1;
    constexpr int N3 = N2 * N1;
    using L = StageLds<N1>;
    __shared__ double lds[L::TOTAL];
    double *s_u = lds + L::U;    // [field][k][j][i]
    double *s_f = lds + L::F;    // [face 0..3][comp 0..5][a + n1 b]
    double *s_D = lds + L::Dm;   // D(i,l) at [i + n1 l]

    const int e = remap_element(blockIdx.x, a.nelt, a.per_xcd);
    if (e >= a.nelt) return;   // whole workgroup leaves together
    const int tid = threadIdx.x;
    const bool act = tid < N2;
    const int i = tid % N1, j = tid / N1;
    const int64_t eo = (int64_t)e * N3;
    const int64_t fs = a.fs;

    // ---- phase 0: element -> LDS (coalesced: lanes cover one k-slice)
    if (act) {
        s_D[tid] = a.D[tid];
#pragma unroll
        for (int f = 0; f < 6; ++f)
#pragma unroll
            for (int k = 0; k < N1; ++k) s_u[f * N3 + k * N2 + tid] = a.uin[f * fs + eo + k * N2 + tid];
    }
    const double *geo = a.geo + (int64_t)e * 9;
    __syncthreads();

    // ---- phase 1: x/y faces with lanes re-mapped onto face node (a,b)=(i,j)
    if (act) {
        double cf[4], um[6], up[6], FH[3], FE[3];
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            const int nb = a.nbr[e * 6 + f];
            int64_t own, nbn;
            if (f == 0) { own = 0 + N1 * i + N2 * j;          nbn = (N1 - 1) + N1 * i + N2 * j; }
            if (f == 1) { own = (N1 - 1) + N1 * i + N2 * j;   nbn = 0 + N1 * i + N2 * j; }
            if (f == 2) { own = i + N2 * j;                   nbn = i + N1 * (N1 - 1) + N2 * j; }
            if (f == 3) { own = i + N1 * (N1 - 1) + N2 * j;   nbn = i + N2 * j; }
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                um[c] = s_u[c * N3 + own];
                up[c] = (nb >= 0) ? a.uin[c * fs + (int64_t)nb * N3 + nbn] : 0.0;
            }
            load_coef<N1, MAT>(a, e, f, tid, cf);
            if (f < 2) face_flux<0>(f == 0 ? -1.0 : 1.0, nb, um, up, cf, geo[3 + f], FH, FE);
            else       face_flux<1>(f == 2 ? -1.0 : 1.0, nb, um, up, cf, geo[3 + f], FH, FE);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                s_f[(f * 6 + c) * N2 + tid] = FH[c];
                s_f[(f * 6 + 3 + c) * N2 + tid] = FE[c];
            }
        }
    }
    __syncthreads();
    if (!act) return;

    // ---- phase 2: one level at a time: volume curl, z faces, lift of the
    // x/y fluxes, inverse mass, RK update
    const double gx = geo[0], gy = geo[1], gz = geo[2];
    const int nbzm = a.nbr[e * 6 + 4], nbzp = a.nbr[e * 6 + 5];
    double ie = 0.0, im = 0.0;
    if (MAT == MAT_ELEM) { ie = a.ieps[e]; im = a.imu[e]; }
#pragma unroll
    for (int k = 0; k < N1; ++k) {
        double Hy_r = 0, Hz_r = 0, Ey_r = 0, Ez_r = 0;
        double Hx_s = 0, Hz_s = 0, Ex_s = 0, Ez_s = 0;
        double Hx_t = 0, Hy_t = 0, Ex_t = 0, Ey_t = 0;
        const double *row = s_u + k * N2 + N1 * j;   // u(l, j, k)
        const double *col = s_u + k * N2 + i;        // u(i, l, k)
        const double *clm = s_u + N1 * j + i;        // u(i, j, l)
#pragma unroll
        for (int l = 0; l < N1; ++l) {
            const double dil = s_D[i + N1 * l], djl = s_D[j + N1 * l], dkl = s_D[k + N1 * l];
            Hy_r += dil * row[1 * N3 + l];
            Hz_r += dil * row[2 * N3 + l];
            Ey_r += dil * row[4 * N3 + l];
            Ez_r += dil * row[5 * N3 + l];
            Hx_s += djl * col[0 * N3 + N1 * l];
            Hz_s += djl * col[2 * N3 + N1 * l];
            Ex_s += djl * col[3 * N3 + N1 * l];
            Ez_s += djl * col[5 * N3 + N1 * l];
            Hx_t += dkl * clm[0 * N3 + N2 * l];
            Hy_t += dkl * clm[1 * N3 + N2 * l];
            Ex_t += dkl * clm[3 * N3 + N2 * l];
            Ey_t += dkl * clm[4 * N3 + N2 * l];
        }
        // resH = -W curl E, resE = W curl H (diagonal metric), per unit J w3
        double r[6];
        r[0] = -(gy * Ez_s - gz * Ey_t);
        r[1] = -(gz * Ex_t - gx * Ez_r);
        r[2] = -(gx * Ey_r - gy * Ex_s);
        r[3] = gy * Hz_s - gz * Hy_t;
        r[4] = gz * Hx_t - gx * Hz_r;
        r[5] = gx * Hy_r - gy * Hx_s;
        const int64_t vnode = k * N2 + tid;
        if (k == 0 || k == N1 - 1) {
            const int zf = (k == 0) ? 4 : 5;
            const int nb = (k == 0) ? nbzm : nbzp;
            const int64_t nbn = (k == 0) ? (N1 - 1) * N2 + tid : tid;
            double cf[4], um[6], up[6], FH[3], FE[3];
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                um[c] = s_u[c * N3 + vnode];
                up[c] = (nb >= 0) ? a.uin[c * fs + (int64_t)nb * N3 + nbn] : 0.0;
            }
            load_coef<N1, MAT>(a, e, zf, tid, cf);
            face_flux<2>(k == 0 ? -1.0 : 1.0, nb, um, up, cf, geo[3 + zf], FH, FE);
#pragma unroll
            for (int c = 0; c < 3; ++c) { r[c] += FH[c]; r[3 + c] += FE[c]; }
        }
        if (i == 0) {
#pragma unroll
            for (int c = 0; c < 6; ++c) r[c] += s_f[(0 * 6 + c) * N2 + j + N1 * k];
        }
        if (i == N1 - 1) {
#pragma unroll
            for (int c = 0; c < 6; ++c) r[c] += s_f[(1 * 6 + c) * N2 + j + N1 * k];
        }
        if (j == 0) {
#pragma unroll
            for (int c = 0; c < 6; ++c) r[c] += s_f[(2 * 6 + c) * N2 + i + N1 * k];
        }
        if (j == N1 - 1) {
#pragma unroll
            for (int c = 0; c < 6; ++c) r[c] += s_f[(3 * 6 + c) * N2 + i + N1 * k];
        }
        // inverse mass: hbm1 = 1/(mu J w3), ebm1 = 1/(eps J w3); J w3 is
        // already divided out of the volume and lift terms
        if (MAT == MAT_NODE) { ie = a.ieps[eo + vnode]; im = a.imu[eo + vnode]; }
#pragma unroll
        for (int c = 0; c < 3; ++c) { r[c] *= im; r[3 + c] *= ie; }
        if (MODE == MODE_RESIDUAL) {
#pragma unroll
            for (int c = 0; c < 6; ++c) a.res[c * fs + eo + vnode] = r[c];
        } else {
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const int64_t p = c * fs + eo + vnode;
                const double kk = a.rka * a.kbuf[p] + a.dt * r[c];
                a.kbuf[p] = kk;
                a.uout[p] = s_u[c * N3 + vnode] + a.rkb * kk;
            }
        }
    }
}

}  // namespace nkc
