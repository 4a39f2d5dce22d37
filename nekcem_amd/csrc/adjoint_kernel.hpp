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
    const double *__restrict__ ieps;
    const double *__restrict__ imu;
    const double *__restrict__ yz;      // (Y, Z) = (sqrt(eps/mu), sqrt(mu/eps)) pairs, like ieps
    int64_t fs;
    int32_t nelt;
    int32_t per_xcd;
    double rka, rkb, dt, C0;
    // Drude region (cem_maxwell_drude): per element offset into the compact
    // per-element blocks (-1: none), per-node (a, b); the adjoint current
    // and its RK register (in place), the stage-input current, and the
    // per-node gradients with respect to a and b (accumulated)
    const int32_t *__restrict__ drude_off;
    const double *__restrict__ dpar;
    double *__restrict__ lamJ;
    double *__restrict__ kapJ;
    const double *__restrict__ stJ;
    double *__restrict__ gA;
    double *__restrict__ gB;
};

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

// Flux coefficients of one side of a face, (Y02, C02Y, Z02, C02Z) =
// (-Y1/(2Y0), C0/(2Y0), Z1/(2Z0), C0/(2Z0)), from hy = 1/(2Y0) and
// hz = 1/(2Z0); the same products the host tables hold.
__device__ __forceinline__ void side_coef(double hy, double hz, double Y1, double Z1, double C0, double c[4]) {
    c[0] = -hy * Y1;
    c[1] = hy * C0;
    c[2] = hz * Z1;
    c[3] = hz * C0;
}

// Gradient of <a, F(coef)> with respect to (Y0, Y1, Z0, Z1) through the
// coefficients above (d(1/(2Y0))/dY0 = -2 hy^2).
__device__ __forceinline__ void coef_bar(double gY02, double gC02Y, double gZ02, double gC02Z,
                                         double hy, double hz, double Y1, double Z1, double C0,
                                         double &Yb0, double &Yb1, double &Zb0, double &Zb1) {
    Yb0 = (gY02 * Y1 - gC02Y * C0) * (2.0 * hy * hy);
    Yb1 = -gY02 * hy;
    Zb0 = -(gZ02 * Z1 + gC02Z * C0) * (2.0 * hz * hz);
    Zb1 = gZ02 * hz;
}

// Transposed flux at one face node of this element (normal sg e_AX), both
// halves of the shared face, plus the eps/mu gradient at the own node.
//   vo, uo: own v = P kt and own stage state (H then E)
//   vn, un: the neighbour's at the paired node (interior faces)
//   L, Ln: own / neighbour lift factors
//   mo: own material; Yn, Zn: the neighbour's admittance and impedance
// The flux coefficients are formed here from the admittances/impedances as
// cem_maxwell_materials averages them (src/cem_maxwell.F:262-325): interior
// Y0 = (Y- + Y+)/2, Y1 = Y+; PEC Y0 = Y1 = Y-; no partner Y0 = Y-/2, Y1 = 0.
// Out: adj (H then E) added to B^T v at the own node; ge, gm gradients.
template <int AX>
__device__ __forceinline__ void adj_face(double sg, int type, const double vo[6], const double uo[6],
                                         const double vn[6], const double un[6],
                                         double L, double Ln, double C0,
                                         const NodeMat &mo, double Yn, double Zn,
                                         double adj[6], double &ge, double &gm) {
    const double Yo = mo.Y, Zo = mo.Z;
    double Ys, Zs;                       // twice the averages
    if (type >= 0) { Ys = Yo + Yn; Zs = Zo + Zn; }
    else if (type == FACE_PEC) { Ys = 2.0 * Yo; Zs = 2.0 * Zo; }
    else { Ys = Yo; Zs = Zo; }
    const double hy = 0.5 / (0.5 * Ys), hz = 0.5 / (0.5 * Zs);
    double cf[4];
    side_coef(hy, hz, Ys - Yo, Zs - Zo, C0, cf);
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
        coef_bar(gY02, gC02Y, gZ02, gC02Z, hy, hz, Ys - Yo, Zs - Zo, C0, Yb0, Yb1, Zb0, Zb1);
        dY = 0.5 * Yb0;
        dZ = 0.5 * Zb0;
        // neighbour half: its normal is -n, its jump vectors equal ours
        double cfn[4];
        side_coef(hy, hz, Ys - Yn, Zs - Zn, C0, cfn);
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
        coef_bar(hY02, hC02Y, hZ02, hC02Z, hy, hz, Ys - Yn, Zs - Zn, C0, Yb0, Yb1, Zb0, Zb1);
        dY += 0.5 * Yb0 + Yb1;
        dZ += 0.5 * Zb0 + Zb1;
    } else if (type == FACE_PEC) {
        double Yb0, Yb1, Zb0, Zb1;
        coef_bar(gY02, gC02Y, gZ02, gC02Z, hy, hz, Yo, Zo, C0, Yb0, Yb1, Zb0, Zb1);
        dY = Yb0 + Yb1;
        dZ = Zb0 + Zb1;
    } else {
        double Yb0, Yb1, Zb0, Zb1;
        coef_bar(gY02, gC02Y, gZ02, gC02Z, hy, hz, 0.0, 0.0, C0, Yb0, Yb1, Zb0, Zb1);
        dY = 0.5 * Yb0;
        dZ = 0.5 * Zb0;
    }
    // Y = sqrt(eps/mu), Z = sqrt(mu/eps): dY/deps = Y/(2 eps), dZ/deps = -Z/(2 eps)
    const double dq = 0.5 * (dY * Yo - dZ * Zo);
    ge = dq * mo.ie;
    gm = -dq * mo.im;
}

// GM = 1: also the eps/mu gradients (needs the stage residual); GM = 0: the
// adjoint state and the Drude gradients only (no residual store is read).
// A 4-wave workgroup is held to 3 waves per SIMD (168 VGPRs, a few spills:
// three elements per CU, as its LDS allows), so that lambda can stay in
// registers from its load to the update instead of being read twice.
template <int N1, int MAT, int GM>
__global__ void __launch_bounds__(Shape<N1>::BS) __attribute__((amdgpu_waves_per_eu(Shape<N1>::BS <= 256 ? 3 : 1)))
adjoint_kernel(AdjArgs a) {
    using S = Shape<N1>;
    constexpr int N2 = S::N2, N3 = S::N3, KH = S::KH, LANES = S::LANES;
    // level loop unrolled except at N1 = 10, where unrolling costs
    // a wave per SIMD (240 VGPRs against 179)
    constexpr int LVL_UNROLL = (N1 >= 10) ? 1 : KH;
    __shared__ double s_v[6 * N3];      // v = P kt, swizzled like s_u
    __shared__ double s_f[32 * N2];     // [face 0..3][adjH 3, adjE 3, geps, gmu][q]
    __shared__ double s_D[N2];
    __shared__ double s_g[GM ? 2 * N3 : 1];   // [grp][k N2 + q]: gradient without the face terms

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

    // ---- phase 0: this role's fields and levels.  v = P kt goes to LDS;
    // the kappa output and the inverse-mass part of the material gradient
    // (d r/d m = -r/m, so -kt.r/m) need no other lane and are finished
    // here, so that the role's lambda, kappa, r and g streams are all
    // loaded up front ----
    if (t < N2) s_D[t] = a.D[t];
    double lm[3][KH];                    // lambda of the role's levels, kept for the update
    if (act) {
        double kp[3][KH], rr[3][KH], gg[KH], pm[KH];
        double *g = grp ? a.geps : a.gmu;
#pragma unroll
        for (int kk = 0; kk < KH; ++kk) {
            const int k = k0 + kk;
            const bool ok = k < N1;
            const int64_t node = eo + k * N2 + q;
            if (MAT == MAT_ELEM) pm[kk] = grp ? a.ieps[e] : a.imu[e];
            else pm[kk] = ok ? (grp ? a.ieps[node] : a.imu[node]) : 0.0;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int64_t p = (fo + c) * fs + node;
                lm[c][kk] = ok ? a.lam_in[p] : 0.0;
                kp[c][kk] = ok ? a.kap_in[p] : 0.0;
                rr[c][kk] = (GM && ok) ? ld_stream<nt_streams<N1>()>(a.rs + p) : 0.0;
            }
            gg[kk] = (GM && ok) ? g[node] : 0.0;
        }
#pragma unroll
        for (int kk = 0; kk < KH; ++kk) {
            const int k = k0 + kk;
            if (k >= N1) break;
            const int64_t node = eo + k * N2 + q;
            double kr = 0.0;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double kt = kp[c][kk] + a.rkb * lm[c][kk];
                s_v[(fo + c) * N3 + S::idx(i, j, k)] = pm[kk] * kt;
                kr += kt * rr[c][kk];
                st_stream<nt_streams<N1>()>(a.kap_out + (fo + c) * fs + node, a.rka * kt);
            }
            if (GM) s_g[grp * N3 + k * N2 + q] = gg[kk] - a.dt * kr * pm[kk];
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
        double vo[6], uo[6], vn[6], un[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            vo[c] = s_v[c * N3 + S::idx(vi, vj, vk)];
            uo[c] = GM ? a.us[c * fs + own] : 0.0;    // the jumps feed the eps/mu gradient only
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
                un[c] = GM ? a.us[c * fs + pn] : 0.0;
            }
            Ln = a.geo[(int64_t)nb * 9 + 3 + (r ^ 1)];
        } else {
#pragma unroll
            for (int c = 0; c < 6; ++c) { vn[c] = 0.0; un[c] = 0.0; }
        }
        double adj[6], ge, gm;
        const double sg = (r & 1) ? 1.0 : -1.0;
        if (r < 2) adj_face<0>(sg, nb, vo, uo, vn, un, geo[3 + r], Ln, a.C0, mo, Yn, Zn, adj, ge, gm);
        else       adj_face<1>(sg, nb, vo, uo, vn, un, geo[3 + r], Ln, a.C0, mo, Yn, Zn, adj, ge, gm);
#pragma unroll
        for (int c = 0; c < 6; ++c) s_f[(r * 8 + c) * N2 + q] = adj[c];
        if (GM) {
            s_f[(r * 8 + 6) * N2 + q] = ge;
            s_f[(r * 8 + 7) * N2 + q] = gm;
        }
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
        double vo[6], uo[6], vn[6], un[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            vo[c] = s_v[c * N3 + S::idx(i, j, k)];
            uo[c] = GM ? a.us[c * fs + node] : 0.0;
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
                un[c] = GM ? a.us[c * fs + pn] : 0.0;
            }
            Ln = a.geo[(int64_t)nbz * 9 + 3 + (zface ^ 1)];
        } else {
#pragma unroll
            for (int c = 0; c < 6; ++c) { vn[c] = 0.0; un[c] = 0.0; }
        }
        double adj[6], ge, gm;
        adj_face<2>(half ? 1.0 : -1.0, nbz, vo, uo, vn, un, geo[3 + zface], Ln, a.C0,
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
        if (grp == 1 && a.drude_off) {
            // r_J = -a J + b E, r_E -= J/eps: lambda_E += dt b ktJ,
            // lambda_J += dt (-v_E - a ktJ), kappa_J = a_i ktJ;
            // d/da = -ktJ.J, d/db = ktJ.E at the stage input
            const int dd = a.drude_off[e];
            if (dd >= 0) {
                const int64_t jb = (int64_t)dd * 3 * N3 + vloc;
                const int64_t pp = (int64_t)dd * N3 + vloc;
                const double pa = a.dpar[pp * 2 + 0], pb = a.dpar[pp * 2 + 1];
                double sJ = 0.0, sE = 0.0;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double lj = a.lamJ[jb + c * N3];
                    const double kt = a.kapJ[jb + c * N3] + a.rkb * lj;
                    const double vE = s_v[(3 + c) * N3 + S::idx(i, j, k)];
                    w[c] += pb * kt;
                    a.lamJ[jb + c * N3] = lj + a.dt * (-vE - pa * kt);
                    a.kapJ[jb + c * N3] = a.rka * kt;
                    sJ += kt * a.stJ[jb + c * N3];
                    sE += kt * a.us[(3 + c) * fs + node];
                }
                a.gA[pp] -= a.dt * sJ;
                a.gB[pp] += a.dt * sE;
            }
        }
        double *g = grp ? a.geps : a.gmu;
        if (GM) g[node] = s_g[grp * N3 + vloc] + a.dt * gface;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int64_t p = (fo + c) * fs + node;
            st_stream<nt_streams<N1>()>(a.lam_out + p, lm[c][kk] + a.dt * w[c]);
        }
    }
}

}  // namespace nkc
