/*
 * nekcem_hip.h -- C-ABI of libnekcem_hip.so, the MI355X (gfx950) DG-SEM
 * Maxwell time step that drops in for NekCEM's cem_maxwell_op_rk.
 *
 * Conventions (NekCEM's, src/jl/name.h:32-41 and SIZE):
 *   - double = Fortran REAL with -fdefault-real-8; index arrays are int32 and
 *     1-based, exactly as the Fortran COMMON arrays hold them;
 *   - volume index i + n1*j + n1^2*k + n1^3*e (column-major), face index
 *     a + n1*b + n1^2*(ef-1) + 6*n1^2*e with ef in Ed's (preprocessor) order
 *     1: s=-1, 2: r=+1, 3: s=+1, 4: r=-1, 5: t=-1, 6: t=+1;
 *   - sizes are int64 (the reference's CUDA prototype used 32-bit sizes,
 *     src/cem_mxm_gpu.cu:34-83);
 *   - every entry point returns an int32 status, NKC_OK = 0; the message of
 *     the last failure on the calling thread is nkc_last_error();
 *   - the per-node (and per-Drude-node) arrays of nkc_set_state /
 *     nkc_get_state / nkc_residual / nkc_curl / nkc_set_materials /
 *     nkc_set_drude / nkc_cem_error_fields / nkc_adjoint_backward /
 *     nkc_adjoint_get_drude / nkc_adjoint_apply / nkc_batch_set_state /
 *     nkc_batch_get_state may be host or device memory of the handle's GPU
 *     (copies use hipMemcpyDefault; tables built on the host stage device
 *     arrays through host memory first); they are synchronous.
 * Fortran binds these through iso_c_binding (nekcem_amd/fortran/
 * nekcem_hip_mod.f90); Python through ctypes (nekcem_amd/_lib.py).
 */
#ifndef NEKCEM_HIP_H
#define NEKCEM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NKC_OK 0
#define NKC_ERR_ARG 1          /* bad handle, pointer or size */
#define NKC_ERR_HIP 2          /* HIP runtime error */
#define NKC_ERR_UNSUPPORTED 3  /* mesh/material form the kernels do not cover */
#define NKC_ERR_BLOWUP 4       /* non-finite or > 1e20 state (chk_blowup) */
#define NKC_ERR_NODEVICE 5     /* no HIP device visible */

typedef struct nkc_handle nkc_handle;

/* Everything cem_maxwell_init leaves in COMMON for the RK path, i.e. the
 * device working set of acc_copy_all_in (src/cem_drive.F:397-566).
 * Array sizes: npts = (N+1)^3*nelt, nxzfl = 6*(N+1)^2*nelt. */
typedef struct nkc_desc {
    int32_t N;                  /* polynomial order (SIZE lxi); lx1 = N+1 */
    int32_t ifupwind;           /* 1: upwind flux (C0=1), 0: central (C0=0),
                                   src/cem_maxwell.F:939-940 */
    int64_t nelt;               /* elements on this GPU */
    const double *dxm1;         /* (N+1)^2 derivative matrix, /DXYZ/ */
    const double *wxm1;         /* (N+1) GLL weights, /WXYZ/ */
    const double *rxmn, *rymn, *rzmn; /* npts cofactor metrics, /minputv/ */
    const double *sxmn, *symn, *szmn;
    const double *txmn, *tymn, *tzmn;
    const double *jacm;         /* npts Jacobian */
    const double *bmn;          /* npts mass jac*w3 */
    const double *unxm, *unym, *unzm; /* nxzfl unit normals, /minputf/ */
    const double *aream;        /* nxzfl face area weights */
    const int32_t *cemface;     /* nxzfl, 1-based volume index, /cemfce1/ */
    const int32_t *face_partner;/* nxzfl, 1-based index of the face node that
                                   shares this node's gsh_face id, 0 = none.
                                   Replaces the gs handle gsh_face
                                   (src/nek5_connect11.F:2179-2233). */
    const int32_t *cempec;      /* ncempec 1-based PEC face-node list, /bdry1/ */
    int64_t ncempec;
    const double *permittivity; /* npts, /materialv/ */
    const double *permeability; /* npts */
    const double *Z_0, *Z_1, *Y_0, *Y_1; /* nxzfl, /materialf/ after
                                   cem_maxwell_materials (:262-325) */
} nkc_desc;

/* A box of nel[0] x nel[1] x nel[2] straight-sided hexahedra on
 * [lo, hi] (element e = ex + nel0*(ey + nel1*ez)), periodic or PEC per axis,
 * with per-element permittivity/permeability: the .box description genbox2
 * reads (src/nek5_genbox.F:2-145), for callers that have no reference arrays.
 * It produces the same tables nkc_create derives from the arrays of the same
 * mesh. */
typedef struct nkc_box_desc {
    int32_t N;                  /* polynomial order */
    int32_t ifupwind;           /* 1 upwind, 0 central */
    int32_t nel[3];             /* elements per axis */
    int32_t pec[3];             /* per axis: 0 periodic, 1 PEC walls */
    double lo[3], hi[3];        /* box extent */
    const double *eps_elem;     /* nelt permittivities, NULL = 1 */
    const double *mu_elem;      /* nelt permeabilities, NULL = 1 */
} nkc_box_desc;

/* Upload the working set and build the on-device face-exchange tables.
 * Mirror of acc_copy_all_in (src/cem_drive.F:397) plus the setup half of
 * cem_maxwell_init (src/cem_maxwell.F:65-191).  Returns NKC_ERR_UNSUPPORTED
 * for geometry other than axis-aligned affine hexahedra (box meshes). */
int32_t nkc_create(const nkc_desc *d, nkc_handle **h);

/* Same handle from a box description (no reference arrays needed). */
int32_t nkc_create_box(const nkc_box_desc *d, nkc_handle **h);

/* A handle from the mesh arrays NekCEM's readers fill (rdmesh/rdbdry and
 * bin_rd1_mesh/bin_rd1_bc, src/nek5_connect2.F:1774-2240): corners
 * xc/yc/zc[8*nelt] (corners 1-4 on t = -1 counter-clockwise from
 * (r,s) = (-1,-1), 5-8 above them), per face in Nek order (1: s=-1, 2: r=+1,
 * 3: s=+1, 4: r=-1, 5: t=-1, 6: t=+1) cbc[(6*e + f)*3 .. +2] (3 characters)
 * and bc[(6*e + f)*5 .. +4].  'E' and 'P' faces name their neighbour element
 * and face in bc parameters 1-2 (1-based), or 0 to find an 'E' neighbour by
 * position (a .re2 lists boundary faces only); 'PEC' is a perfect
 * conductor.  Elements must be axis-aligned boxes; materials per element
 * (NULL: 1).  Node coordinates for the on-device diagnostics are set with
 * nkc_set_coords. */
int32_t nkc_create_hex(int32_t N, int32_t ifupwind, int64_t nelt, const double *xc,
                       const double *yc, const double *zc, const char *cbc, const double *bc,
                       const double *eps_elem, const double *mu_elem, nkc_handle **h);

/* Mirror of acc_copy_all_out (src/cem_drive.F:482): frees the handle. */
int32_t nkc_destroy(nkc_handle *h);

/* !$ACC UPDATE DEVICE / HOST of hn, en, khn, ken (src/EMWAVE:4-23).  Each is
 * 3 components at stride ld (= lpts1 in the reference) with ld >= npts;
 * entries npts..ld-1 of each component are neither read nor written.
 * khn/ken may be NULL: set_state then zeroes k (a1 = 0 makes k from a
 * previous step irrelevant), get_state skips them.  After nkc_step, khn/ken
 * hold the register of the step's fourth stage unless nkc_set_keep_k(h, 1)
 * is in effect (see there). */
int32_t nkc_set_state(nkc_handle *h, const double *hn, const double *en,
                      const double *khn, const double *ken, int64_t ld);
int32_t nkc_get_state(nkc_handle *h, double *hn, double *en, double *khn,
                      double *ken, int64_t ld);

/* nsteps calls of cem_maxwell_op_rk (src/cem_maxwell.F:327-345): 5-stage
 * low-storage RK per step, each stage one fused kernel, the step replayed
 * from a captured hipGraph.  *time += nsteps*dt (time_advancing_pde,
 * src/cem_drive.F:639).  Asynchronous on the handle's stream; the next
 * get_state or nkc_sync waits for it. */
int32_t nkc_step(nkc_handle *h, double *time, double dt, int64_t nsteps);

/* The RK register khn/ken as rk4_upd leaves it (src/cem_common.F:18-76).
 * rk4_upd writes k in all five stages, but the next step's first stage
 * multiplies it by a_1 = 0, so by default (keep = 0) nkc_step's last stage
 * skips that write (10 % of a step's bytes): khn/ken read after a step then
 * hold the fourth stage's register, while the fields are unaffected.  With
 * keep = 1 the last stage writes it, and khn/ken after nkc_step (and
 * nkc_adjoint_forward) equal the reference's.  The Drude, Lorentz, PML and
 * graphene registers are always written. */
int32_t nkc_set_keep_k(nkc_handle *h, int32_t keep);

/* One evaluation of cem_maxwell_op (src/cem_maxwell.F:484-508) on the
 * current state: reshn/resen = inverse-mass-scaled residual, 3 components at
 * stride ld.  Test entry point for the stage composition. */
int32_t nkc_residual(nkc_handle *h, double rktime, double *reshn, double *resen,
                     int64_t ld);

/* The volume term alone on the current state: cem_maxwell's 3-D branch
 * (src/cem_maxwell.F:574-582, maxwell_wght_curl :1428-1497, then chsign):
 * resen = W curl H, reshn = -W curl E with W = bm1, before the flux and the
 * inverse mass.  Test entry point for the weighted curl. */
int32_t nkc_curl(nkc_handle *h, double *reshn, double *resen, int64_t ld);

/* usersrc-style source on E (the Gaussian pulse of the configs): every
 * stage adds amp(x) * dir * exp(-((t - t0)/tw)^2) to resE before the inverse
 * mass, with amp (npts) bm1-weighted as a usersrc hook would add it
 * (src/cem_maxwell.F:500-504) and t the stage time rktime.  amp = NULL
 * removes the source. */
int32_t nkc_set_source(nkc_handle *h, const double *amp, const double dir[3],
                       double t0, double tw);

/* The source's time factor: kind 0 the Gaussian pulse of nkc_set_source
 * (which restores it), kind 1 g(t) = sin(omega t + phase) -- the dipole of
 * tests/3dboxpml/3dboxpml.usr:30-88 (tfac = sin(-omega tt): omega < 0,
 * phase 0). */
int32_t nkc_set_source_time(nkc_handle *h, int32_t kind, double omega, double phase);

/* PML (src/cem_maxwell_pml.F).  The stage adds pml_step
 * (src/cem_maxwell_pml.F:508-592) after the flux and before usersrc, and
 * advances the auxiliary fields B = pmlbn and D = pmldn by the same RK stage
 * (rk_maxwell_ab, src/cem_maxwell.F:1942-1954).  Faces with the boundary
 * code 'PML' are walls (cem_maxwell_flux_pec treats them as PEC): build the
 * mesh with them as PEC.
 * nkc_set_pml takes the reference's own arrays: pmlptr (maxpml element
 * numbers, 1-based, ascending), pmlsigma(lpts1,3) and the initial pmlbn,
 * pmldn (3 components at stride ld each; NULL for zero).  maxpml = 0
 * removes the PML.
 * nkc_setup_pml builds them as cem_maxwell_init does (pml_fill_faceary,
 * march_faces, pml_extent_and_tags, pml_calc_sigma; src/cem_maxwell_pml.F:
 * 85-505; thick, order, referr = param(77..79)) from pml_face[6 nelt], 1
 * for the faces whose code is 'PML' (reference face order 1..6 per
 * element): the layers are the `thick` elements inward from those faces,
 * sigma follows from the node coordinates (box extents or nkc_set_coords)
 * and the current eps and mu; B = D = 0.  extents (12 or NULL) receives
 * pmlinner[6] then pmlouter[6] (order x-, x+, y-, y+, z-, z+).
 * nkc_set_pml_fields sets B and D (as userini does) and zeroes their RK
 * register.  nkc_get_pml copies back (which = 0) B then D, (1) their RK
 * register, (2) the last residual from nkc_residual (respmlbn, respmldn),
 * each 6 components at stride ld, or (3) sigma (3 components), on the PML
 * elements' nodes only.  The adjoint refuses a PML. */
int32_t nkc_set_pml(nkc_handle *h, int64_t maxpml, const int32_t *pmlptr, const double *sigma,
                    const double *bn, const double *dn, int64_t ld);
int32_t nkc_setup_pml(nkc_handle *h, const int8_t *pml_face, int32_t thick, double order,
                      double referr, double *extents);
int32_t nkc_set_pml_fields(nkc_handle *h, const double *bn, const double *dn, int64_t ld);
int32_t nkc_get_pml(nkc_handle *h, int32_t which, double *out, int64_t ld);

/* Incident plane wave, the device form of userinc
 * (tests/3ddielectric/3ddielectric.usr:6-47, called by cem_maxwell_op,
 * src/cem_maxwell.F:497-498): on each listed face the face traces get
 *   H += hamp u,  E += eamp u,  u = cos(k . x - omega t)
 * before the flux, x the trace's node and t the stage time, and the
 * neighbour across the face sees the same traces.  faces holds nface pairs
 * (element 1-based, face 1..6 in the reference's face order).  Needs node
 * coordinates (box extents or nkc_set_coords).  nface = 0 removes it.  The
 * adjoint refuses it. */
int32_t nkc_set_incident(nkc_handle *h, int64_t nface, const int32_t *faces, const double k[3],
                         double omega, const double hamp[3], const double eamp[3]);

/* Graphene sheet current, the device form of userfsrc +
 * cem_3d_graphene_current (tests/3dgraphene/3dgraphene.usr:235-262,
 * src/cem_maxwell.F:2827-2932, called from cem_maxwell_flux3d before the
 * face gather, :941-942): at each of the ngraph face points graphindex
 * (1-based, NekCEM's face-point numbering (e-1) 6 nxzf + (f-1) nxzf + i)
 * the side's sheet current J is formed from its face traces (the incident
 * wave included) and its state, subtracted from -n x H before the gather,
 * and the state (fjn slots 2..6: a Drude term and two critical-point pairs)
 * advances with the stage.  graphparams(ld, 12) holds the twelve sheet
 * parameters per point (a_d, b_d, b_cp1, a_211, a_221, b_11, b_21, b_cp2,
 * a_212, a_222, b_12, b_22), fjn(ld, 3, 6) the initial state in slots
 * 2..6 (slot 1, the current, is formed each stage and not read); the RK
 * register starts at zero.  Yfac = 0.5/Y_0 and yconduc come from the
 * current materials and follow nkc_set_materials.  Both sides of a sheet
 * are listed, as uservp does.  ngraph = 0 removes it.  nkc_residual
 * includes the current of the present state without advancing it; the
 * adjoint refuses a sheet. */
int32_t nkc_set_graphene(nkc_handle *h, int64_t ngraph, const int32_t *graphindex,
                         const double *graphparams, const double *fjn, int64_t ld);

/* Copy back the sheet state (which = 0) or its RK register (1) into slots
 * 2..6 of fjn(ld, 3, 6) at the listed points; other entries are left as
 * they are. */
int32_t nkc_get_graphene(nkc_handle *h, int32_t which, double *fjn, int64_t ld);

/* Drude ADE of cem_maxwell_drude (src/cem_maxwell.F:3095-3147) over the
 * ndrude volume nodes dindex (1-based): resE -= J*bmn, resJ = -a J + b E,
 * J advanced by the same low-storage RK.  params is (a, b) per listed node
 * (a = gamma, b = omega_p^2), interleaved [2*l + 0/1]; J and kJ are
 * [3][ndrude] (component-major) or NULL for zero.  ndrude = 0 removes it. */
int32_t nkc_set_drude(nkc_handle *h, int64_t ndrude, const int32_t *dindex,
                      const double *params, const double *J, const double *kJ);

/* Copy the Drude current, its RK register and the last residual (from
 * nkc_residual) back, [3][ndrude] each; any pointer may be NULL. */
int32_t nkc_get_drude(nkc_handle *h, double *J, double *kJ, double *resJ);

/* Lorentz ADE (cem_maxwell_lorentz, src/cem_maxwell.F:3149-3211) on nlor
 * volume nodes lindex (1-based): resE -= J*bmn, resJ = -a J - b P + c E,
 * resP = J, J and P advanced by the same low-storage RK, i.e. eps(w) =
 * eps_oo + c/(b - i a w - w^2).  params is (a, b, c) per listed node,
 * interleaved [3*l + 0/1/2]; L and kL are [6][nlor] (Jx, Jy, Jz, Px, Py,
 * Pz; component-major) or NULL for zero.  nlor = 0 removes it.  The adjoint
 * does not cover it (nkc_adjoint_forward returns NKC_ERR_UNSUPPORTED). */
int32_t nkc_set_lorentz(nkc_handle *h, int64_t nlor, const int32_t *lindex,
                        const double *params, const double *L, const double *kL);
/* Copy L, its RK register and the last residual back ([6][nlor] each;
 * any pointer may be NULL). */
int32_t nkc_get_lorentz(nkc_handle *h, double *L, double *kL, double *resL);

/* Replace the materials by per-node permittivity and permeability (npts
 * each): 1/eps, 1/mu and every face node's flux coefficients are rebuilt as
 * cem_maxwell_materials does (src/cem_maxwell.F:262-325). */
int32_t nkc_set_materials(nkc_handle *h, const double *eps, const double *mu);

/* ---- discrete adjoint (no counterpart in the reference) ----------------
 * nkc_adjoint_forward advances nsteps RK steps from the current state like
 * nkc_step, keeping a checkpoint of the state every `every` steps.
 * nkc_adjoint_backward takes dL/dH, dL/dE at the final state (3 components
 * at stride ld) and returns dL/deps and dL/dmu per node (npts, may be NULL):
 * segments are recomputed from their checkpoints keeping every stage, then
 * the RK stages are reversed (transposed curl, transposed flux/lift,
 * reversed low-storage RK).  The forward's final state, RK registers and
 * time are restored.  With a Drude region (nkc_set_drude) its current is
 * checkpointed and reversed too, and nkc_adjoint_get_drude returns the
 * gradients with respect to the per-node a (gamma) and b (omega_p^2) in the
 * order of the listed nodes; the loss is taken not to depend on J(T).
 * nkc_set_materials, nkc_set_source and nkc_set_drude discard a stored
 * forward. */
int32_t nkc_adjoint_forward(nkc_handle *h, double *time, double dt, int64_t nsteps,
                            int64_t every);
int32_t nkc_adjoint_backward(nkc_handle *h, const double *lamH, const double *lamE,
                             int64_t ld, double *geps, double *gmu);
int32_t nkc_adjoint_get_drude(nkc_handle *h, double *ga, double *gb);
/* A segment is reversed in windows of S <= every steps whose 5 S stages (state
 * and residual, plus the Drude current) are kept at once; window starts are
 * recomputed from the segment checkpoint.  S is the largest that fits in
 * 70 % of the free device memory, or at most max_steps when that is > 0
 * (0 restores the default); it applies from the next nkc_adjoint_forward.
 * When whole segments fit (S = every, no cap), the forward keeps the stages
 * of its last K segments, K as many as half of the free memory holds, and
 * the backward sweep recomputes only the others.  So by default the adjoint
 * may hold up to about half of the device's free memory from
 * nkc_adjoint_forward until nkc_adjoint_release, the next forward, a
 * set_* call or nkc_destroy; a cap (max_steps > 0) holds it to the
 * checkpoints plus one window of 5 max_steps stages.
 * nkc_adjoint_info: out[5] = {every, segments, S, K, device bytes held}. */
int32_t nkc_adjoint_set_store(nkc_handle *h, int64_t max_steps);
/* Free everything the adjoint holds (checkpoints, stage store, adjoint
 * state, gradients); a backward sweep then needs a new forward. */
int32_t nkc_adjoint_release(nkc_handle *h);
/* material = 1 (default): the backward sweep returns dL/deps and dL/dmu; 0:
 * it skips them (zeros are returned), which drops the stage residual from
 * the store (half its size: larger windows) and from the adjoint stage;
 * the Drude gradients are unaffected.  Applies from the next
 * nkc_adjoint_forward. */
int32_t nkc_adjoint_set_grads(nkc_handle *h, int32_t material);
int32_t nkc_adjoint_info(nkc_handle *h, int64_t *out);
/* aH, aE = B^T P (kH, kE) on the current state: the transpose of the
 * operator whose residual is nkc_residual (without sources).  Test entry. */
int32_t nkc_adjoint_apply(nkc_handle *h, const double *kH, const double *kE, double *aH,
                          double *aE, int64_t ld);

/* ---- on-device diagnostics (userchk without a host copy) ---------------
 * Analytic solutions of the reference's test cases, evaluated on the device:
 *   NKC_SOL_BOXPER: tests/3dboxper/3dboxper.usr:45-86, periodic (1,1,1)
 *     mode, omega = sqrt(3);
 *   NKC_SOL_BOXPEC: tests/3dboxpec/3dboxpec.usr:91-113, PEC cavity mode,
 *     omega = pi sqrt(3).
 * nkc_set_coords gives the node coordinates (xm1, ym1, zm1, npts each);
 * box handles (nkc_create_box) have them already.  nkc_init_solution sets
 * the state to the solution at `time` (userini; k := 0).  nkc_cem_error is
 * cem_error (src/cem_common.F:1335-1355) of the state against the solution
 * at `time`, nkc_cem_error_fields against caller-given exact fields (hn/en
 * layout, stride ld): out[12] = mass-weighted L2 of Hx, Hy, Hz, Ex, Ey, Ez,
 * then their Linf. */
#define NKC_SOL_BOXPER 1
#define NKC_SOL_BOXPEC 2
int32_t nkc_set_coords(nkc_handle *h, const double *xm1, const double *ym1, const double *zm1);
int32_t nkc_init_solution(nkc_handle *h, int32_t sol, double time);
int32_t nkc_cem_error(nkc_handle *h, int32_t sol, double time, double *out);
int32_t nkc_cem_error_fields(nkc_handle *h, const double *hx, const double *ex, int64_t ld,
                             double *out);

/* nkc_cem_error without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same kernels and the same out[12] as nkc_cem_error, bit for
 * bit; the partial norms go to a fixed 107 KB device array of the library
 * (one per GPU, held until the process ends) instead of handle memory.
 * The node coordinates must already be on the device (nkc_set_coords, or
 * on a box handle one earlier nkc_cem_error / nkc_init_solution), else
 * NKC_ERR_ARG; a mesh needing more partial norms than the array holds
 * (never, at its 1024-workgroup cap) is refused with NKC_ERR_ARG.  The
 * array is shared by every call on a GPU: calls that overlap on one GPU
 * (from several threads or streams) race. */
int32_t nkc_cem_error_noalloc(nkc_handle *h, int32_t sol, double time, double *out);

/* ---- restart files ---------------------------------------------------
 * nkc_restart_write writes the current state as cem_restart_out does
 * (src/io.F:411-490 with src/io_co.c:115-554, one rank): a big-endian
 * legacy-VTK binary unstructured grid -- the header line "istep dumpno nx1
 * np time dt" (time and dt with six decimals, as %f prints them), the node
 * coordinates as float32, the (N)^3 GLL sub-hexahedra of each element, then
 * E and H as VECTORS of doubles (single = 0, param(87) = 0) or floats.
 * Needs the node coordinates (box handles have them; others
 * nkc_set_coords).  nkc_restart_read is restart_swap (src/io.F:637-788):
 * E and H of a file written for this mesh become the state (k := 0) and
 * the header's istep, dumpno, time and dt come back (restart_swap adds the
 * time read to the run's time).  A file of another order or point count
 * is refused. */
int32_t nkc_restart_write(nkc_handle *h, const char *path, int32_t istep, int32_t dumpno, double time, double dt,
                          int32_t single);
int32_t nkc_restart_read(nkc_handle *h, const char *path, int32_t *istep, int32_t *dumpno, double *time,
                         double *dt);

/* chk_blowup (src/cem_common.F:1150-1165): *maxabs = max |u| over the six
 * fields; returns NKC_ERR_BLOWUP if a value is non-finite or above 1e20. */
int32_t nkc_check_finite(nkc_handle *h, double *maxabs);

/* Wait for all work queued on the handle's stream. */
int32_t nkc_sync(nkc_handle *h);

/* The HIP stream (hipStream_t) every kernel of this handle runs on, so a
 * caller can time it with events or order its own work after it. */
int32_t nkc_stream(nkc_handle *h, void **stream);

/* Facts about the compiled handle: out[0] = npts, out[1] = material mode
 * (0 per element, 1 per node), out[2] = elements per workgroup,
 * out[3] = threads per workgroup, out[4] = LDS bytes per workgroup,
 * out[5] = graph capture enabled (1/0). */
int32_t nkc_info(nkc_handle *h, int64_t out[8]);

/* Enable (1, default) or disable (0) hipGraph replay in nkc_step. */
int32_t nkc_set_graph(nkc_handle *h, int32_t enable);

/* Stage-kernel layout (no counterpart in the reference).  Layouts 0..3
 * take one element per workgroup.  Bit 0: the four lane roles of an
 * element get whole waves each (0) or their (N+1)^2 lanes packed back to
 * back in fewer waves (1).  Bit 1: the register allocator is left free (0)
 * or held to 128 VGPRs for 4 waves per SIMD (1).  Layouts 4 and 5 take two
 * elements per workgroup, packed, with the free / 128-VGPR register budget;
 * they are built for N <= 5 only.  Layouts 6 and 7 split an element six
 * ways (the two update groups x three runs of levels, packed; every lane
 * forms one face flux), with the free / occupancy-bound register budget;
 * they are built for N >= 7 and carry sources, Drude and Lorentz regions
 * but no PML, incident field or graphene sheet (nkc_set_layout refuses
 * them on such a handle with NKC_ERR_UNSUPPORTED, and the tuner does not
 * time them there).  The first nkc_step times the layouts on
 * the handle's mesh and keeps the fastest (an autotuner; a packed
 * one-element layout with the same workgroup as the unpacked one, and a
 * layout not built for the order, are not timed).  nkc_set_layout forces
 * 0..NKC_NLAYOUT-1 (one built for the order), or -1 to tune again at the
 * next step; nkc_tune tunes now: out[NKC_NLAYOUT + 1] = {us per stage for
 * each layout (NaN: not timed), chosen layout} (out may be NULL).  All
 * layouts compute the same arithmetic: results are bit-identical. */
#define NKC_NLAYOUT 8
int32_t nkc_set_layout(nkc_handle *h, int32_t layout);
int32_t nkc_tune(nkc_handle *h, double *out);
/* The current layout and the last tuner times without tuning:
 * out[NKC_NLAYOUT + 1] as for nkc_tune (layout -1 before the first step). */
int32_t nkc_get_layout(nkc_handle *h, double *out);

/* ---- batches of independent cavities --------------------------------
 * Many independent box cavities of different sizes and polynomial orders
 * advance together: cavities of one order share one element table, each
 * order's RK stage is one launch, the orders run side by side, and a whole
 * step is one graph replay.  The reference's only way to run independent
 * cases at once is one MPI session per case (neknek, src/nek5_comm_mpi.F:
 * 45-119); no entry point there corresponds to these. */
typedef struct nkc_batch nkc_batch;
int32_t nkc_batch_create(int32_t ncav, const nkc_box_desc *d, nkc_batch **b);
int32_t nkc_batch_destroy(nkc_batch *b);
/* State of cavity cav (0-based): hn/en 3 components at stride ld; k := 0. */
int32_t nkc_batch_set_state(nkc_batch *b, int32_t cav, const double *hn,
                            const double *en, int64_t ld);
int32_t nkc_batch_get_state(nkc_batch *b, int32_t cav, double *hn, double *en,
                            int64_t ld);
/* nsteps RK steps of every cavity, cavity c with its own dt[c]; times[c]
 * (may be NULL) receives each cavity's time afterwards. */
int32_t nkc_batch_step(nkc_batch *b, const double *dt, double *times, int64_t nsteps);
int32_t nkc_batch_sync(nkc_batch *b);
int32_t nkc_batch_stream(nkc_batch *b, void **stream);
/* out[0] = total points, out[1] = cavities, out[2] = polynomial orders
 * (launches per stage), out[3] = total elements. */
int32_t nkc_batch_info(nkc_batch *b, int64_t out[8]);
/* Per polynomial order of the batch (tuned at its first step, -1 before):
 * out[2 i] = N, out[2 i + 1] = stage-kernel layout; cap = ints in out. */
int32_t nkc_batch_layout(nkc_batch *b, int32_t *out, int64_t cap);

/* Message of the last failed call on this thread ("" if none). */
const char *nkc_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* NEKCEM_HIP_H */
