// stage6_kernel.hpp -- the fused RK stage at high orders (N = 7, 8, 9) with
// six lane roles, for gfx950.
//
// The same stage as stage_kernel.hpp (cem_maxwell_op + rk_maxwell_ab,
// src/cem_maxwell.F:327-345, 484-508, 922-1002, 1428-1497, 1833-1904;
// rk4_upd, src/cem_common.F:18-76) and the same arithmetic, term for term
// and in the same order, so its results are bit-identical to every other
// layout's.  What differs is the shape.  With four roles of (N+1)^2 lanes
// each, N = 8 fills 81 of every 128 lanes (two waves per role) and the two
// level halves take 5 and 4 levels; the face phase gives two roles a
// second (z) face.  Here the element is split six ways:
//
//   role r = 2 part + grp: grp 0 updates H from -curl E, grp 1 E from
//   curl H, over the levels of part (N+1 levels in three runs, the longer
//   runs first: 3+3+3 at N = 8, 4+3+3 at N = 9, 3+3+2 at N = 7);
//   the six roles' (N+1)^2 lanes are packed back to back (N = 8: 486 of
//   512 lanes busy, 8 waves);
//   in the face phase role r forms the flux of face r (kernel order x-,
//   x+, y-, y+, z-, z+) at node (i, j), so every lane has exactly one face
//   and the face normal is a run-time unit vector (no divergent branch in
//   the waves that straddle two roles).
//
// Every global load of the element is issued up front (one HBM round
// trip); the face fluxes go through LDS (6 faces x 6 components x
// (N+1)^2) to the lanes of the nodes they lift into.  Physics: sources,
// Drude and Lorentz regions (FEAT = 1); the PML, the incident field and
// the graphene sheet stay with the four-role kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "stage_kernel.hpp"

namespace nkc {

template <int N1>
struct Shape6 {
    static constexpr int N2 = N1 * N1, N3 = N2 * N1;
    static constexpr int ROLES = 6;
    static constexpr int BS = ((ROLES * N2 + 63) / 64) * 64;
    static constexpr int WAVES = BS / 64;
    static constexpr int BASE = N1 / 3, REM = N1 % 3;
    static constexpr int KP = BASE + (REM > 0 ? 1 : 0);   // levels of the longest run
    static constexpr int LDS_BYTES = 8 * (6 * N3 + 36 * N2 + N2);
    // workgroups per CU that the LDS allows (160 KB), and the waves per SIMD
    // they need: the register budget of the OCC instance
    static constexpr int WG_LDS = (160 * 1024) / LDS_BYTES;
    static constexpr int OCC_WAVES = (WG_LDS * WAVES + 3) / 4 > 8 ? 8 : (WG_LDS * WAVES + 3) / 4;
    __device__ static __forceinline__ int k0(int p) { return p * BASE + (p < REM ? p : REM); }
    __device__ static __forceinline__ int kn(int p) { return BASE + (p < REM ? 1 : 0); }
    __device__ static __forceinline__ int idx(int i, int j, int k) { return Shape<N1>::idx(i, j, k); }
};

// Flux at one face node for the face whose outward normal is the unit
// vector n = sg e_AX given as a vector: the same terms as face_flux<AX>
// (every product with a component of n is exact, so the values agree bit
// for bit, the sign of a zero aside).
__device__ __forceinline__ void face_flux_n(const double n[3], int type, const double um[6],
                                            const double up[6], const double cf[4], double L,
                                            double FH[3], double FE[3]) {
    double sE[3], sH[3];
    if (type >= 0) {
        const double dE[3] = {up[3] - um[3], up[4] - um[4], up[5] - um[5]};
        const double dH[3] = {up[0] - um[0], up[1] - um[1], up[2] - um[2]};
        gcross(n, dE, sE);
        gcross(n, dH, sH);
    } else {
        const double Em[3] = {-um[3], -um[4], -um[5]};
        gcross(n, Em, sE);
        if (type == FACE_PEC) {
            sE[0] *= 2.0; sE[1] *= 2.0; sE[2] *= 2.0;
            sH[0] = 0.0; sH[1] = 0.0; sH[2] = 0.0;
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
        FH[c] = L * (cf[0] * sE[c] - cf[1] * fu[c]);
        FE[c] = L * (cf[2] * sH[c] - cf[3] * fw[c]);
    }
}

// OCC = 1 holds the registers to what the LDS-bound number of workgroups
// per CU needs (Shape6::OCC_WAVES waves per SIMD); OCC = 0 leaves the
// allocator free.
template <int N1, int MAT, int MODE, int OCC, int FEAT>
__global__ void __launch_bounds__(Shape6<N1>::BS) __attribute__((amdgpu_waves_per_eu(OCC ? Shape6<N1>::OCC_WAVES : 1)))
stage6_kernel(StageArgs a) {
    static_assert(FEAT <= 1, "the six-role kernel carries sources, Drude and Lorentz only");
    using S = Shape6<N1>;
    constexpr int N2 = S::N2, N3 = S::N3, KP = S::KP, BS = S::BS;
    constexpr bool NT = nt_streams<N1>();
    __shared__ double s_u[6 * N3];
    __shared__ double s_f[36 * N2];     // [face][comp 0..5][a + n1 b]
    __shared__ double s_D[N2];

    const int e = stage_element<(FEAT >= 1)>(a, blockIdx.x, a.nelt);
    if (e >= a.nelt) return;
    const int t = threadIdx.x;
    const int role = t / N2, q = t % N2;
    const bool act = role < S::ROLES;
    const int i = q % N1, j = q / N1;
    const int grp = role & 1, part = role >> 1;
    const int k0 = act ? S::k0(part) : 0, kn = act ? S::kn(part) : 0;
    const int64_t eo = (int64_t)e * N3;
    const int64_t fs = a.fs;
    const int f = role;                      // this lane's face in phase 1
    const int nbf = act ? a.nbr[e * 6 + f] : FACE_NONE;

    // ---- phase 0: every global load of the element, issued up front ----
    constexpr int NLOAD = (6 * N3 + BS - 1) / BS;
    double ld[NLOAD];
#pragma unroll
    for (int n = 0; n < NLOAD; ++n) {
        const int v = t + BS * n;
        ld[n] = (v < 6 * N3) ? a.uin[(v / N3) * fs + eo + (v % N3)] : 0.0;
    }
    constexpr bool KPRE = !OCC;
    // KPF (the occupancy-bound budget without in-stage physics): the first
    // level's register up front, each next one a level ahead
    constexpr bool KPF = !KPRE && FEAT == 0;
    double kp[3][KPRE ? KP : 1];
    if (MODE == MODE_RK && act) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int kk = 0; kk < (KPRE ? KP : 1); ++kk) {
                const int k = k0 + kk;
                kp[c][kk] = ((KPRE || KPF) && (a.kflags & 1) && kk < kn)
                                ? ld_stream<NT>(a.kbuf + (3 * grp + c) * fs + eo + k * N2 + q) : 0.0;
            }
    }
    const int so = (FEAT && grp == 1 && a.src_off && act) ? a.src_off[e] : -1;
    double samp[KP];
#pragma unroll
    for (int kk = 0; kk < KP; ++kk)
        samp[kk] = (so >= 0 && kk < kn) ? a.src_amp[(int64_t)so * N3 + (k0 + kk) * N2 + q] : 0.0;
    // this lane's face: own node, the neighbour's node on the opposite face
    int own_f = 0;
    int64_t nbn = 0;
    switch (f) {
    case 0: own_f = S::idx(0, i, j);      nbn = (N1 - 1) + N1 * i + N2 * j; break;
    case 1: own_f = S::idx(N1 - 1, i, j); nbn = N1 * i + N2 * j; break;
    case 2: own_f = S::idx(i, 0, j);      nbn = i + N1 * (N1 - 1) + N2 * j; break;
    case 3: own_f = S::idx(i, N1 - 1, j); nbn = i + N2 * j; break;
    case 4: own_f = S::idx(i, j, 0);      nbn = i + N1 * j + N2 * (N1 - 1); break;
    default: own_f = S::idx(i, j, N1 - 1); nbn = i + N1 * j; break;
    }
    double up[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) up[c] = (act && nbf >= 0) ? a.uin[c * fs + (int64_t)nbf * N3 + nbn] : 0.0;

    if (t < N2) s_D[t] = a.D[t];
#pragma unroll
    for (int n = 0; n < NLOAD; ++n) {
        const int v = t + BS * n;
        if (v < 6 * N3) {
            const int fld = v / N3, r = v % N3;
            s_u[fld * N3 + S::idx(r % N1, (r / N1) % N1, r / N2)] = ld[n];
        }
    }
    const double *geo = a.geo + (int64_t)e * 9;
    __syncthreads();

    // ---- phase 1: role r forms the flux of face r at node (i, j) ----
    if (act) {
        double um[6], cf[4], FH[3], FE[3];
#pragma unroll
        for (int c = 0; c < 6; ++c) um[c] = s_u[c * N3 + own_f];
        load_coef<N1, MAT>(a, e, f, q, cf);
        const double sg = (f & 1) ? 1.0 : -1.0;
        const int ax = f >> 1;
        const double nv[3] = {ax == 0 ? sg : 0.0, ax == 1 ? sg : 0.0, ax == 2 ? sg : 0.0};
        face_flux_n(nv, nbf, um, up, cf, geo[3 + f], FH, FE);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            s_f[(f * 6 + c) * N2 + q] = FH[c];
            s_f[(f * 6 + 3 + c) * N2 + q] = FE[c];
        }
    }
    __syncthreads();
    if (!act) return;

    // ---- phase 2: this role's levels ----
    const double gx = geo[0], gy = geo[1], gz = geo[2];
    const int src = grp ? 0 : 3;
    const double sgn = grp ? 1.0 : -1.0;
    const double *S0 = s_u + (src + 0) * N3;
    const double *S1 = s_u + (src + 1) * N3;
    const double *S2 = s_u + (src + 2) * N3;
    double minv = 0.0;
    if (MAT == MAT_ELEM) minv = grp ? a.ieps[e] : a.imu[e];
    const double dt = a.dt_e ? a.dt_e[e] : a.dt;
    double sgt = 0.0;
    if (so >= 0) {
        const double tstage = *a.tdev + a.rkc * dt;
        if (a.src_kind == 1) {
            sgt = sin(a.src_om * tstage + a.src_ph);
        } else {
            const double tt = (tstage - a.src_t0) / a.src_tw;
            sgt = exp(-tt * tt);
        }
    }
    const int dd = (FEAT && grp == 1 && a.drude_off) ? a.drude_off[e] : -1;
    const int lo = (FEAT && grp == 1 && a.lor_off) ? a.lor_off[e] : -1;
    const int co = 3 * grp;
    // the x face (0 x-, 1 x+) and the y face (2 y-, 3 y+) this lane's nodes
    // lie on, or -1, and the LDS offsets of their fluxes (level 0)
    const int xface = (i == 0) ? 0 : (i == N1 - 1 ? 1 : -1);
    const int yface = (j == 0) ? 2 : (j == N1 - 1 ? 3 : -1);
    const int xbase = ((xface < 0 ? 0 : xface) * 6 + 3 * grp) * N2 + j;
    const int ybase = ((yface < 0 ? 2 : yface) * 6 + 3 * grp) * N2 + i;
    // the plain N = 8 instance holds this lane's r row of D in registers
    // (fits its 128-VGPR budget): N+1 fewer LDS reads per level, which
    // bound the level loop at this order
    constexpr bool DREGR = N1 == 9 && FEAT == 0;
    constexpr bool DREGS = DREGR && OCC;
    double Dr[DREGR ? N1 : 1], Ds[DREGS ? N1 : 1];
    if (DREGR) {
#pragma unroll
        for (int l = 0; l < N1; ++l) Dr[l] = s_D[i + N1 * l];
    }
    if (DREGS) {
#pragma unroll
        for (int l = 0; l < N1; ++l) Ds[l] = s_D[j + N1 * l];
    }
#pragma unroll
    for (int kk = 0; kk < KP; ++kk) {
        if (kk >= kn) break;
        const int k = k0 + kk;
        const int vloc = k * N2 + q;
        double kold[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) kold[c] = kp[c][KPRE ? kk : 0];
        if (KPF && MODE == MODE_RK && (a.kflags & 1) && kk + 1 < kn) {
            // the next level's RK register, issued before this level's
            // stores (a load issued after them is waited on together with
            // them: one in-order counter)
#pragma unroll
            for (int c = 0; c < 3; ++c) kp[c][0] = ld_stream<NT>(a.kbuf + (3 * grp + c) * fs + eo + vloc + N2);
        } else if (!KPRE && !KPF && MODE == MODE_RK && (a.kflags & 1)) {
#pragma unroll
            for (int c = 0; c < 3; ++c) kold[c] = ld_stream<NT>(a.kbuf + (3 * grp + c) * fs + eo + vloc);
        }
        double S1_r = 0, S2_r = 0, S0_s = 0, S2_s = 0, S0_t = 0, S1_t = 0;
#pragma unroll
        for (int l = 0; l < N1; ++l) {
            const int pr = S::idx(l, j, k), ps = S::idx(i, l, k), pt = S::idx(i, j, l);
            const double dkl = s_D[k + N1 * l];
            const double drl = DREGR ? Dr[DREGR ? l : 0] : s_D[i + N1 * l];
            const double dsl = DREGS ? Ds[DREGS ? l : 0] : s_D[j + N1 * l];
            S1_r += drl * S1[pr];
            S2_r += drl * S2[pr];
            S0_s += dsl * S0[ps];
            S2_s += dsl * S2[ps];
            S0_t += dkl * S0[pt];
            S1_t += dkl * S1[pt];
        }
        double r[3];
        r[0] = sgn * (gy * S2_s - gz * S1_t);
        r[1] = sgn * (gz * S0_t - gx * S2_r);
        r[2] = sgn * (gx * S1_r - gy * S0_s);
        const int own = S::idx(i, j, k);
        // lifted fluxes: the z face first, then x-, x+, y-, y+ (the order
        // of every other layout)
        if (k == 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) r[c] += s_f[(4 * 6 + co + c) * N2 + q];
        }
        if (k == N1 - 1) {
#pragma unroll
            for (int c = 0; c < 3; ++c) r[c] += s_f[(5 * 6 + co + c) * N2 + q];
        }
        // a node lies on at most one x face and one y face: one lift each,
        // x before y as in the face order
        if (xface >= 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) r[c] += s_f[xbase + c * N2 + N1 * k];
        }
        if (yface >= 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) r[c] += s_f[ybase + c * N2 + N1 * k];
        }
        const int64_t node = eo + k * N2 + q;
        if (so >= 0) {
            const double am = samp[kk] * sgt;
#pragma unroll
            for (int c = 0; c < 3; ++c) r[c] += am * a.src_dir[c];
        }
        if (dd >= 0) {
            const int64_t jb = (int64_t)dd * 3 * N3 + vloc;
            const double pa = a.dpar[((int64_t)dd * N3 + vloc) * 2 + 0];
            const double pb = a.dpar[((int64_t)dd * N3 + vloc) * 2 + 1];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double Jc = a.dJ[jb + c * N3];
                const double rJ = -pa * Jc + pb * s_u[(3 + c) * N3 + own];
                r[c] -= Jc;
                if (MODE == MODE_RESIDUAL) {
                    a.dresJ[jb + c * N3] = rJ;
                } else {
                    if (a.st_J) a.st_J[jb + c * N3] = Jc;
                    const double kj = a.rka * a.dkJ[jb + c * N3] + dt * rJ;
                    a.dkJ[jb + c * N3] = kj;
                    a.dJ[jb + c * N3] = Jc + a.rkb * kj;
                }
            }
        }
        if (lo >= 0) {
            const int64_t lb = (int64_t)lo * 6 * N3 + vloc;
            const double *lp = a.lpar + ((int64_t)lo * N3 + vloc) * 3;
            const double la = lp[0], lbp = lp[1], lc = lp[2];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double Jc = a.lL[lb + c * N3], Pc = a.lL[lb + (3 + c) * N3];
                const double rJ = -la * Jc - lbp * Pc + lc * s_u[(3 + c) * N3 + own];
                r[c] -= Jc;
                if (MODE == MODE_RESIDUAL) {
                    a.lresL[lb + c * N3] = rJ;
                    a.lresL[lb + (3 + c) * N3] = Jc;
                } else {
                    const double kj = a.rka * a.lkL[lb + c * N3] + dt * rJ;
                    const double kp2 = a.rka * a.lkL[lb + (3 + c) * N3] + dt * Jc;
                    a.lkL[lb + c * N3] = kj;
                    a.lkL[lb + (3 + c) * N3] = kp2;
                    a.lL[lb + c * N3] = Jc + a.rkb * kj;
                    a.lL[lb + (3 + c) * N3] = Pc + a.rkb * kp2;
                }
            }
        }
        if (MAT == MAT_NODE) minv = grp ? a.ieps[node] : a.imu[node];
#pragma unroll
        for (int c = 0; c < 3; ++c) r[c] *= minv;
        if (MODE == MODE_RESIDUAL) {
#pragma unroll
            for (int c = 0; c < 3; ++c) st_stream<NT>(a.res + (co + c) * fs + node, r[c]);
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int64_t p = (co + c) * fs + node;
                const double kk2 = a.rka * kold[c] + dt * r[c];
                if (a.kflags & 2) st_stream<NT>(a.kbuf + p, kk2);
                st_stream<NT>(a.uout + p, s_u[(co + c) * N3 + own] + a.rkb * kk2);
            }
            if (a.st_r) {
#pragma unroll
                for (int c = 0; c < 3; ++c) st_stream<NT>(a.st_r + (co + c) * fs + node, r[c]);
            }
        }
    }
}

}  // namespace nkc
