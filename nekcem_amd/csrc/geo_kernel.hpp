// geo_kernel.hpp -- the RK stage on general hexahedra (curved, rotated,
// sheared elements; any conforming face pairing), for gfx950.
//
// The same stage as stage_kernel.hpp (cem_maxwell_op + rk_maxwell_ab,
// src/cem_maxwell.F:484-508, 1906-1968), with the geometry per node instead
// of per element: maxwell_wght_curl's full cofactor combine
// (src/cem_maxwell.F:1464-1495) in strong form, curl u = sum_a G_a x du/dr_a
// with G_ab = (cofactor ab)/J, the flux of cem_maxwell_flux3d
// (:922-1002) with each face node's own unit normal and partner (any
// in-face orientation), the lift area/(J w3) of add_flux_to_res and
// invqmass (:654-756, 1833-1904).
//
// Mapping: one workgroup of 256 lanes per element; the element's six fields
// are staged in LDS; phase 1 forms the flux of every face node (6 (N+1)^2
// of them, strided over the lanes) into LDS, phase 2 gives each lane whole
// nodes (strided): nine derivatives of E and of H along r, s, t, the two
// curls, the lifted fluxes of the faces the node lies on, inverse mass and
// the low-storage RK update.  Up to N = 11 (12^3 nodes: 125 KB of LDS, one
// workgroup per CU).  Materials per node, flux coefficients per face node.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "stage_kernel.hpp"

namespace nkc {

enum GeoMode : int { GMODE_RK = 0, GMODE_RESIDUAL = 1, GMODE_CURL = 2 };

struct GeoArgs {
    const double *__restrict__ uin;    // 6 fields, stride fs
    double *__restrict__ uout;         // GMODE_RK
    double *__restrict__ kbuf;         // GMODE_RK: RK register in place
    double *__restrict__ res;          // GMODE_RESIDUAL / GMODE_CURL
    const double *__restrict__ D;      // D(i,l) at D[i + n1 l]
    const double *__restrict__ gmet;   // 9 x npts metric quotients (component major)
    const double *__restrict__ bmw;    // npts: bm1 = J w3 (GMODE_CURL)
    const double *__restrict__ fgeo;   // 4 per face node: n_x, n_y, n_z, area/bm1
    const int32_t *__restrict__ fnb;   // per face node: partner volume node / FaceType
    const double *__restrict__ fco;    // 4 per face node: Y02, C02Y, Z02, C02Z
    const double *__restrict__ ieps;   // per node
    const double *__restrict__ imu;
    int64_t fs;
    int64_t npts;
    int32_t nelt;
    int32_t per_xcd;
    double rka, rkb, dt;
    int32_t kflags;                    // bit 0 read k, bit 1 write k
};

constexpr int GEO_BS = 256;

// In-element volume index of in-face node q = a + n1 b on the reference's
// face ef0 + 1 (0: y-, 1: x+, 2: y+, 3: x-, 4: z-, 5: z+), skpdat order
// (src/nek5_connect11.F:1483-1528).
template <int N1>
__device__ __forceinline__ int ed_face_node(int ef0, int q) {
    constexpr int N2 = N1 * N1, M = N1 - 1;
    const int a = q % N1, b = q / N1;
    switch (ef0) {
    case 0: return a + N2 * b;
    case 1: return M + N1 * a + N2 * b;
    case 2: return a + N1 * M + N2 * b;
    case 3: return N1 * a + N2 * b;
    case 4: return a + N1 * b;
    default: return a + N1 * b + N2 * M;
    }
}

template <int N1, int MODE>
__global__ void __launch_bounds__(GEO_BS) geo_stage_kernel(GeoArgs a) {
    constexpr int N2 = N1 * N1, N3 = N2 * N1;
    __shared__ double s_u[6 * N3];
    __shared__ double s_f[6 * N2 * 6];   // [face node of the element][FH 0..2, FE 0..2]
    __shared__ double s_D[N2];
    const int e = remap_element(blockIdx.x, a.nelt, a.per_xcd);
    if (e >= a.nelt) return;
    const int t = threadIdx.x;
    const int64_t eo = (int64_t)e * N3, fs = a.fs;
    const int64_t so = (int64_t)e * 6 * N2;

    for (int v = t; v < 6 * N3; v += GEO_BS) s_u[v] = a.uin[(v / N3) * fs + eo + v % N3];
    for (int v = t; v < N2; v += GEO_BS) s_D[v] = a.D[v];
    __syncthreads();

    // ---- phase 1: the flux at every face node of the element ----
    if (MODE != GMODE_CURL) {
        for (int fq = t; fq < 6 * N2; fq += GEO_BS) {
            const int ef0 = fq / N2, q = fq % N2;
            const int64_t s = so + fq;
            const int own = ed_face_node<N1>(ef0, q);
            const double n[3] = {a.fgeo[4 * s], a.fgeo[4 * s + 1], a.fgeo[4 * s + 2]};
            const double L = a.fgeo[4 * s + 3];
            const double cf[4] = {a.fco[4 * s], a.fco[4 * s + 1], a.fco[4 * s + 2], a.fco[4 * s + 3]};
            const int32_t nb = a.fnb[s];
            double um[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) um[c] = s_u[c * N3 + own];
            double sE[3], sH[3];
            if (nb >= 0) {
                // n x [E] = (-n x E-) + (-n' x E+), n' = -n
                double dE[3], dH[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    dH[c] = a.uin[c * fs + nb] - um[c];
                    dE[c] = a.uin[(3 + c) * fs + nb] - um[3 + c];
                }
                gcross(n, dE, sE);
                gcross(n, dH, sH);
            } else {
                // no partner: gs leaves -n x E-, -n x H-; PEC doubles E, zeroes H
                const double Em[3] = {-um[3], -um[4], -um[5]};
                gcross(n, Em, sE);
                if (nb == FACE_PEC) {
                    sE[0] *= 2.0; sE[1] *= 2.0; sE[2] *= 2.0;
                    sH[0] = sH[1] = sH[2] = 0.0;
                } else {
                    const double Hm[3] = {-um[0], -um[1], -um[2]};
                    gcross(n, Hm, sH);
                }
            }
            double fu[3], fw[3];
            gcross(n, sH, fu);
            gcross(n, sE, fw);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                s_f[fq * 6 + c] = L * (cf[0] * sE[c] - cf[1] * fu[c]);
                s_f[fq * 6 + 3 + c] = L * (cf[2] * sH[c] - cf[3] * fw[c]);
            }
        }
        __syncthreads();
    }

    // ---- phase 2: whole nodes per lane ----
    for (int v = t; v < N3; v += GEO_BS) {
        const int i = v % N1, j = (v / N1) % N1, k = v / N2;
        // d/dr, d/ds, d/dt of the six fields
        double dr[6], ds[6], dt_[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) { dr[c] = 0.0; ds[c] = 0.0; dt_[c] = 0.0; }
        for (int l = 0; l < N1; ++l) {
            const double di = s_D[i + N1 * l], dj = s_D[j + N1 * l], dk = s_D[k + N1 * l];
            const int pr = l + N1 * j + N2 * k, ps = i + N1 * l + N2 * k, pt = i + N1 * j + N2 * l;
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                dr[c] += di * s_u[c * N3 + pr];
                ds[c] += dj * s_u[c * N3 + ps];
                dt_[c] += dk * s_u[c * N3 + pt];
            }
        }
        const int64_t p = eo + v;
        double G[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) G[c] = a.gmet[(int64_t)c * a.npts + p];
        // gradient component b of field c: dr G[0 + b] + ds G[3 + b] + dt G[6 + b]
        auto dx = [&](int c, int b) { return dr[c] * G[b] + ds[c] * G[3 + b] + dt_[c] * G[6 + b]; };
        // rH = -curl E, rE = curl H
        double r[6];
        r[0] = -(dx(5, 1) - dx(4, 2));
        r[1] = -(dx(3, 2) - dx(5, 0));
        r[2] = -(dx(4, 0) - dx(3, 1));
        r[3] = dx(2, 1) - dx(1, 2);
        r[4] = dx(0, 2) - dx(2, 0);
        r[5] = dx(1, 0) - dx(0, 1);
        if (MODE == GMODE_CURL) {
            const double w = a.bmw[p];
#pragma unroll
            for (int c = 0; c < 6; ++c) a.res[c * fs + p] = w * r[c];
            continue;
        }
        // lifted fluxes of the faces through this node (reference face order)
        auto lift = [&](int ef0, int q) {
#pragma unroll
            for (int c = 0; c < 6; ++c) r[c] += s_f[(ef0 * N2 + q) * 6 + c];
        };
        if (j == 0) lift(0, i + N1 * k);
        if (i == N1 - 1) lift(1, j + N1 * k);
        if (j == N1 - 1) lift(2, i + N1 * k);
        if (i == 0) lift(3, j + N1 * k);
        if (k == 0) lift(4, i + N1 * j);
        if (k == N1 - 1) lift(5, i + N1 * j);
        const double im = a.imu[p], ie = a.ieps[p];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            r[c] *= im;
            r[3 + c] *= ie;
        }
        if (MODE == GMODE_RESIDUAL) {
#pragma unroll
            for (int c = 0; c < 6; ++c) a.res[c * fs + p] = r[c];
        } else {
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const int64_t x = c * fs + p;
                const double kold = (a.kflags & 1) ? a.kbuf[x] : 0.0;
                const double kn = a.rka * kold + a.dt * r[c];
                if (a.kflags & 2) a.kbuf[x] = kn;
                a.uout[x] = s_u[c * N3 + v] + a.rkb * kn;
            }
        }
    }
}

}  // namespace nkc
