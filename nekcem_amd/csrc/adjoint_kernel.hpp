// adjoint_kernel.hpp -- one reversed RK stage of the discrete adjoint.
//
// Forward stage i (stage_kernel.hpp): r = P B(u) + P s, k <- a_i k + dt r,
// u <- u + b_i k, with P = diag(1/mu, 1/eps) per node, B the strong-form DG
// operator (volume curl + lifted upwind/central flux) and s the source.
// Reversing it, with lambda the adjoint of u and kappa the adjoint of k:
//   kt = kappa + b_i lambda,  v = P kt
//   lambda <- lambda + dt B^T v          (B^T: Dt-curl + transposed flux/lift)
//   kappa  <- a_i kt
//   g_eps  += dt (d r/d eps)^T kt,  g_mu += dt (d r/d mu)^T kt
// where r depends on eps/mu through P (d r_E/d eps = -r_E/eps) and through
// the impedance averages of the face coefficients (Y = sqrt(eps/mu),
// Z = sqrt(mu/eps), as cem_maxwell_materials forms them, src/cem_maxwell.F:
// 262-325).  The stage input state u and the stage residual r come from the
// forward recompute.  Each face node gathers both halves of the transposed
// flux of its face: its own flux and the neighbour's flux on the shared face,
// so every output has a single owner (no atomics).
#pragma once
#include "stage_kernel.hpp"

namespace nkc {

struct AdjArgs {
    const double *__restrict__ lam_in;  // 6 fields, stride fs
    const double *__restrict__ kap_in;
    double *__restrict__ lam_out;
    double *__restrict__ kap_out;
    const double *__restrict__ us;      // stage input state
    const double *__restrict__ rs;      // stage residual (inverse-mass scaled)
    double *__restrict__ geps;          // npts, accumulated
    double *__restrict__ gmu;
    const double *__restrict__ D;
    const double *__restrict__ geo;
    const int32_t *__restrict__ nbr;
    const double *__restrict__ fco;
    const double *__restrict__ ieps;
    const double *__restrict__ imu;
    const double *__restrict__ yz;      // (Y, Z) = (sqrt(eps/mu), sqrt(mu/eps)) pairs, like ieps
    int64_t fs;
    int32_t nelt;
    int32_t per_xcd;
    double rka, rkb, dt, C0;
};

template <int N1, int MAT>
__device__ __forceinline__ const double *coef_ptr(const AdjArgs &a, int e, int f, int q) {
    return (MAT == MAT_ELEM) ? a.fco + ((int64_t)e * 6 + f) * 4
                             : a.fco + (((int64_t)e * 6 + f) * (N1 * N1) + q) * 4;
}

// Material data at one node: inverse permittivity and permeability and the
// admittance/impedance pair, from the per-element or per-node tables.
struct NodeMat {
    double ie, im, Y, Z;
};

template <int N1, int MAT>
__device__ __forceinline__ NodeMat node_mat(const AdjArgs &a, int e, int64_t node) {
    const int64_t x = (MAT == MAT_ELEM) ? (int64_t)e : node;
    const double2 yz = *reinterpret_cast<const double2 *>(a.yz + 2 * x);
    return NodeMat{a.ieps[x], a.imu[x], yz.x, yz.y};
}

// Gradient of <a, F(coef)> with respect to (Y0, Y1, Z0, Z1) through
// (Y02, C02Y, Z02, C02Z) = (-Y1/(2Y0), C0/(2Y0), Z1/(2Z0), C0/(2Z0)).
__device__ __forceinline__ void coef_bar(double gY02, double gC02Y, double gZ02, double gC02Z,
                                         double Y0, double Y1, double Z0, double Z1, double C0,
                                         double &Yb0, double &Yb1, double &Zb0, double &Zb1) {
    const double hy = 0.5 / Y0, hz = 0.5 / Z0;
    Yb0 = (gY02 * Y1 - gC02Y * C0) * (hy / Y0);
    Yb1 = -gY02 * hy;
    Zb0 = -(gZ02 * Z1 + gC02Z * C0) * (hz / Z0);
    Zb1 = gZ02 * hz;
}

// Transposed flux at one face node of this element (normal sg e_AX), both
// halves of the shared face, plus the eps/mu gradient at the own node.
//   vo, uo: own v = P kt and own stage state (H then E)
//   vn, un: the neighbour's at the paired node (interior faces)
//   cf, cfn: own / neighbour flux coefficients, L, Ln: lift factors
//   mo: own material; Yn, Zn: the neighbour's admittance and impedance
// Out: adj (H then E) added to B^T v at the own node; ge, gm gradients.
template <int AX>
__device__ __forceinline__ void adj_face(double sg, int type, const double vo[6], const double uo[6],
                                         const double vn[6], const double un[6], const double cf[4],
                                         const double cfn[4], double L, double Ln, double C0,
                                         const NodeMat &mo, double Yn, double Zn,
                                         double adj[6], double &ge, double &gm) {
    double aH[3], aE[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) { aH[c] = L * vo[c]; aE[c] = L * vo[3 + c]; }
    double sE[3], sH[3];
    if (type >= 0) {
        double dE[3] = {un[3] - uo[3], un[4] - uo[4], un[5] - uo[5]};
        double dH[3] = {un[0] - uo[0], un[1] - uo[1], un[2] - uo[2]};
        ncross<AX>(sg, dE, sE);
        ncross<AX>(sg, dH, sH);
    } else {
        double Em[3] = {-uo[3], -uo[4], -uo[5]};
        double Hm[3] = {-uo[0], -uo[1], -uo[2]};
        ncross<AX>(sg, Em, sE);
        if (type == FACE_PEC) {
            sE[0] *= 2.0; sE[1] *= 2.0; sE[2] *= 2.0;
            sH[0] = sH[1] = sH[2] = 0.0;
        } else {
            ncross<AX>(sg, Hm, sH);
        }
    }
    // own half: F_H = Y02 sE - C02Y n x sH, F_E = Z02 sH - C02Z n x sE
    double NaH[3], NaE[3], NNaH[3], NNaE[3];
    ncross<AX>(sg, aH, NaH);
    ncross<AX>(sg, aE, NaE);
    ncross<AX>(sg, NaH, NNaH);
    ncross<AX>(sg, NaE, NNaE);
    if (type == FACE_PEC) {
        // sE = -2 n x E-, sH = 0
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            adj[c] = 0.0;
            adj[3 + c] = 2.0 * cf[0] * NaH[c] + 2.0 * cf[3] * NNaE[c];
        }
    } else {
        // d/d(jump): gE = -Y02 N aH - C02Z N^2 aE ; gH = -C02Y N^2 aH - Z02 N aE;
        // the own traces enter the jumps with a minus sign
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            adj[3 + c] = cf[0] * NaH[c] + cf[3] * NNaE[c];
            adj[c] = cf[1] * NNaH[c] + cf[2] * NaE[c];
        }
    }
    const double Yo = mo.Y, Zo = mo.Z;
    double NsH[3], NsE[3];
    ncross<AX>(sg, sH, NsH);
    ncross<AX>(sg, sE, NsE);
    double gY02 = 0, gC02Y = 0, gZ02 = 0, gC02Z = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        gY02 += aH[c] * sE[c];
        gC02Y -= aH[c] * NsH[c];
        gZ02 += aE[c] * sH[c];
        gC02Z -= aE[c] * NsE[c];
    }
    double dY, dZ;
    if (type >= 0) {
        double Yb0, Yb1, Zb0, Zb1;
        coef_bar(gY02, gC02Y, gZ02, gC02Z, 0.5 * (Yo + Yn), Yn, 0.5 * (Zo + Zn), Zn, C0, Yb0, Yb1, Zb0, Zb1);
        dY = 0.5 * Yb0;
        dZ = 0.5 * Zb0;
        // neighbour half: its normal is -n, its jump vectors equal ours
        double bH[3], bE[3], NbH[3], NbE[3], NNbH[3], NNbE[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) { bH[c] = Ln * vn[c]; bE[c] = Ln * vn[3 + c]; }
        ncross<AX>(sg, bH, NbH);
        ncross<AX>(sg, bE, NbE);
        ncross<AX>(sg, NbH, NNbH);
        ncross<AX>(sg, NbE, NNbE);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            adj[3 + c] += cfn[0] * NbH[c] - cfn[3] * NNbE[c];
            adj[c] += -cfn[1] * NNbH[c] + cfn[2] * NbE[c];
        }
        double hY02 = 0, hC02Y = 0, hZ02 = 0, hC02Z = 0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            hY02 += bH[c] * sE[c];
            hC02Y += bH[c] * NsH[c];
            hZ02 += bE[c] * sH[c];
            hC02Z += bE[c] * NsE[c];
        }
        coef_bar(hY02, hC02Y, hZ02, hC02Z, 0.5 * (Yn + Yo), Yo, 0.5 * (Zn + Zo), Zo, C0, Yb0, Yb1, Zb0, Zb1);
        dY += 0.5 * Yb0 + Yb1;
        dZ += 0.5 * Zb0 + Zb1;
    } else if (type == FACE_PEC) {
        double Yb0, Yb1, Zb0, Zb1;
        coef_bar(gY02, gC02Y, gZ02, gC02Z, Yo, Yo, Zo, Zo, C0, Yb0, Yb1, Zb0, Zb1);
        dY = Yb0 + Yb1;
        dZ = Zb0 + Zb1;
    } else {
        double Yb0, Yb1, Zb0, Zb1;
        coef_bar(gY02, gC02Y, gZ02, gC02Z, 0.5 * Yo, 0.0, 0.5 * Zo, 0.0, C0, Yb0, Yb1, Zb0, Zb1);
        dY = 0.5 * Yb0;
        dZ = 0.5 * Zb0;
    }
    // Y = sqrt(eps/mu), Z = sqrt(mu/eps): dY/deps = Y/(2 eps), dZ/deps = -Z/(2 eps)
    const double dq = 0.5 * (dY * Yo - dZ * Zo);
    ge = dq * mo.ie;
    gm = -dq * mo.im;
}

template <int N1, int MAT>
__global__ void __launch_bounds__(Shape<N1>::BS) NKC_OCC
adjoint_kernel(AdjArgs a) {
    using S = Shape<N1>;
    constexpr int N2 = S::N2, N3 = S::N3, KH = S::KH, LANES = S::LANES;
    // level loop unrolled except at N1 = 8, where the unrolled loop
    // is scheduled into 244 VGPRs (2 waves/SIMD) against 164 rolled
    constexpr int LVL_UNROLL = (N1 == 8) ? 1 : KH;
    __shared__ double s_v[6 * N3];      // v = P kt, swizzled like s_u
    __shared__ double s_f[32 * N2];     // [face 0..3][adjH 3, adjE 3, geps, gmu][q]
    __shared__ double s_D[N2];

    const int e = remap_element(blockIdx.x, a.nelt, a.per_xcd);
    if (e >= a.nelt) return;
    const int t = threadIdx.x;
    const int role = t / LANES, q = t % LANES;
    const bool act = q < N2;
    const int i = q % N1, j = q / N1;
    const int grp = role & 1;            // 0: H fields, 1: E fields
    const int half = role >> 1;
    const int k0 = half ? KH : 0;
    const int64_t eo = (int64_t)e * N3;
    const int64_t fs = a.fs;
    const int fo = 3 * grp;              // this role's field offset

    // ---- phase 0: this role's fields and levels; v = P (kappa + b lambda) into LDS ----
    if (t < N2) s_D[t] = a.D[t];
    if (act) {
#pragma unroll
        for (int kk = 0; kk < KH; ++kk) {
            const int k = k0 + kk;
            if (k >= N1) break;
            const int64_t node = eo + k * N2 + q;
            double pm;
            if (MAT == MAT_ELEM) pm = grp ? a.ieps[e] : a.imu[e];
            else pm = grp ? a.ieps[node] : a.imu[node];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int64_t p = (fo + c) * fs + node;
                s_v[(fo + c) * N3 + S::idx(i, j, k)] = pm * (a.kap_in[p] + a.rkb * a.lam_in[p]);
            }
        }
    }
    const double *geo = a.geo + (int64_t)e * 9;
    __syncthreads();

    // ---- phase 1: role r takes x/y face r at face node (i,j) ----
    if (act) {
        const int r = role;
        const int nb = a.nbr[e * 6 + r];
        int vi, vj, vk;
        int64_t nbn;
        if (r == 0) { vi = 0;      vj = i;      vk = j; nbn = (N1 - 1) + N1 * i + N2 * j; }
        if (r == 1) { vi = N1 - 1; vj = i;      vk = j; nbn = 0 + N1 * i + N2 * j; }
        if (r == 2) { vi = i;      vj = 0;      vk = j; nbn = i + N1 * (N1 - 1) + N2 * j; }
        if (r == 3) { vi = i;      vj = N1 - 1; vk = j; nbn = i + N2 * j; }
        const int64_t own = eo + vi + N1 * vj + N2 * vk;
        double vo[6], uo[6], vn[6], un[6], cfn[4] = {0, 0, 0, 0};
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            vo[c] = s_v[c * N3 + S::idx(vi, vj, vk)];
            uo[c] = a.us[c * fs + own];
        }
        double Yn = 1.0, Zn = 1.0, Ln = 0.0;
        const NodeMat mo = node_mat<N1, MAT>(a, e, own);
        if (nb >= 0) {
            const int64_t pn = (int64_t)nb * N3 + nbn;
            const NodeMat mn = node_mat<N1, MAT>(a, nb, pn);
            Yn = mn.Y;
            Zn = mn.Z;
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const double pm = (c < 3) ? mn.im : mn.ie;
                vn[c] = pm * (a.kap_in[c * fs + pn] + a.rkb * a.lam_in[c * fs + pn]);
                un[c] = a.us[c * fs + pn];
            }
            const double *pc = coef_ptr<N1, MAT>(a, nb, r ^ 1, q);
            cfn[0] = pc[0]; cfn[1] = pc[1]; cfn[2] = pc[2]; cfn[3] = pc[3];
            Ln = a.geo[(int64_t)nb * 9 + 3 + (r ^ 1)];
        } else {
#pragma unroll
            for (int c = 0; c < 6; ++c) { vn[c] = 0.0; un[c] = 0.0; }
        }
        const double *pc = coef_ptr<N1, MAT>(a, e, r, q);
        const double cf[4] = {pc[0], pc[1], pc[2], pc[3]};
        double adj[6], ge, gm;
        const double sg = (r & 1) ? 1.0 : -1.0;
        if (r < 2) adj_face<0>(sg, nb, vo, uo, vn, un, cf, cfn, geo[3 + r], Ln, a.C0, mo, Yn, Zn, adj, ge, gm);
        else       adj_face<1>(sg, nb, vo, uo, vn, un, cf, cfn, geo[3 + r], Ln, a.C0, mo, Yn, Zn, adj, ge, gm);
#pragma unroll
        for (int c = 0; c < 6; ++c) s_f[(r * 8 + c) * N2 + q] = adj[c];
        s_f[(r * 8 + 6) * N2 + q] = ge;
        s_f[(r * 8 + 7) * N2 + q] = gm;
    }
    __syncthreads();
    if (!act) return;

    // ---- phase 2: this role's levels ----
    const double gx = geo[0], gy = geo[1], gz = geo[2];
    const int src = grp ? 0 : 3;          // v of the other group
    const double sgn = grp ? 1.0 : -1.0;
    const double *S0 = s_v + (src + 0) * N3;
    const double *S1 = s_v + (src + 1) * N3;
    const double *S2 = s_v + (src + 2) * N3;
    const int zface = half ? 5 : 4;
    const int nbz = a.nbr[e * 6 + zface];
    // z face of this half (level kz), for this role's fields, ahead of the
    // level loop so that its temporaries are dead before the curl sums
    const int kz = half ? N1 - 1 : 0;
    double zadj[3], zg;
    {
        const int k = kz;
        const int64_t node = eo + k * N2 + q;
        const int64_t nbn = half ? (int64_t)q : (int64_t)(N1 - 1) * N2 + q;
        double vo[6], uo[6], vn[6], un[6], cfn[4] = {0, 0, 0, 0};
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            vo[c] = s_v[c * N3 + S::idx(i, j, k)];
            uo[c] = a.us[c * fs + node];
        }
        double Yn = 1.0, Zn = 1.0, Ln = 0.0;
        const NodeMat mo = node_mat<N1, MAT>(a, e, node);
        if (nbz >= 0) {
            const int64_t pn = (int64_t)nbz * N3 + nbn;
            const NodeMat mn = node_mat<N1, MAT>(a, nbz, pn);
            Yn = mn.Y;
            Zn = mn.Z;
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const double pm = (c < 3) ? mn.im : mn.ie;
                vn[c] = pm * (a.kap_in[c * fs + pn] + a.rkb * a.lam_in[c * fs + pn]);
                un[c] = a.us[c * fs + pn];
            }
            const double *pc = coef_ptr<N1, MAT>(a, nbz, zface ^ 1, q);
            cfn[0] = pc[0]; cfn[1] = pc[1]; cfn[2] = pc[2]; cfn[3] = pc[3];
            Ln = a.geo[(int64_t)nbz * 9 + 3 + (zface ^ 1)];
        } else {
#pragma unroll
            for (int c = 0; c < 6; ++c) { vn[c] = 0.0; un[c] = 0.0; }
        }
        const double *pc = coef_ptr<N1, MAT>(a, e, zface, q);
        const double cf[4] = {pc[0], pc[1], pc[2], pc[3]};
        double adj[6], ge, gm;
        adj_face<2>(half ? 1.0 : -1.0, nbz, vo, uo, vn, un, cf, cfn, geo[3 + zface], Ln, a.C0,
                    mo, Yn, Zn, adj, ge, gm);
#pragma unroll
        for (int c = 0; c < 3; ++c) zadj[c] = adj[fo + c];
        zg = grp ? ge : gm;
    }
    double Dr[N1], Ds[N1];
#pragma unroll
    for (int l = 0; l < N1; ++l) { Dr[l] = s_D[l + N1 * i]; Ds[l] = s_D[l + N1 * j]; }
#pragma unroll LVL_UNROLL
    for (int kk = 0; kk < KH; ++kk) {
        const int k = k0 + kk;
        if (k >= N1) break;
        // transposed volume curl: sums over l with D(l, .)
        double S1_r = 0, S2_r = 0, S0_s = 0, S2_s = 0, S0_t = 0, S1_t = 0;
#pragma unroll
        for (int l = 0; l < N1; ++l) {
            const int pr = S::idx(l, j, k), ps = S::idx(i, l, k), pt = S::idx(i, j, l);
            const double dlk = s_D[l + N1 * k];
            S1_r += Dr[l] * S1[pr];
            S2_r += Dr[l] * S2[pr];
            S0_s += Ds[l] * S0[ps];
            S2_s += Ds[l] * S2[ps];
            S0_t += dlk * S0[pt];
            S1_t += dlk * S1[pt];
        }
        double w[3];
        w[0] = sgn * (gy * S2_s - gz * S1_t);
        w[1] = sgn * (gz * S0_t - gx * S2_r);
        w[2] = sgn * (gx * S1_r - gy * S0_s);
        double gface = 0.0;
        const int vloc = k * N2 + q;
        const int64_t node = eo + vloc;
        if (k == kz) {
#pragma unroll
            for (int c = 0; c < 3; ++c) w[c] += zadj[c];
            gface += zg;
        }
        if (i == 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) w[c] += s_f[(0 * 8 + fo + c) * N2 + j + N1 * k];
            gface += s_f[(0 * 8 + 6 + (1 - grp)) * N2 + j + N1 * k];
        }
        if (i == N1 - 1) {
#pragma unroll
            for (int c = 0; c < 3; ++c) w[c] += s_f[(1 * 8 + fo + c) * N2 + j + N1 * k];
            gface += s_f[(1 * 8 + 6 + (1 - grp)) * N2 + j + N1 * k];
        }
        if (j == 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) w[c] += s_f[(2 * 8 + fo + c) * N2 + i + N1 * k];
            gface += s_f[(2 * 8 + 6 + (1 - grp)) * N2 + i + N1 * k];
        }
        if (j == N1 - 1) {
#pragma unroll
            for (int c = 0; c < 3; ++c) w[c] += s_f[(3 * 8 + fo + c) * N2 + i + N1 * k];
            gface += s_f[(3 * 8 + 6 + (1 - grp)) * N2 + i + N1 * k];
        }
        // inverse-mass part of the material gradient: d r/d m = -r/m; the
        // role's lambda, kappa (re-read, cached from phase 0) and the stage
        // residual
        const double im = (MAT == MAT_ELEM) ? (grp ? a.ieps[e] : a.imu[e])
                                            : (grp ? a.ieps[node] : a.imu[node]);
        double lam[3], kt[3], kr = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int64_t p = (fo + c) * fs + node;
            lam[c] = a.lam_in[p];
            kt[c] = a.kap_in[p] + a.rkb * lam[c];
            kr += kt[c] * ld_stream(a.rs + p);
        }
        const double gmat = gface - kr * im;
        double *g = grp ? a.geps : a.gmu;
        g[node] += a.dt * gmat;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int64_t p = (fo + c) * fs + node;
            st_stream(a.lam_out + p, lam[c] + a.dt * w[c]);
            st_stream(a.kap_out + p, a.rka * kt[c]);
        }
    }
}

}  // namespace nkc
