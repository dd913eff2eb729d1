/* canary_dg.h -- C ABI of the MI355X nodal-DG library (libcanary_dg.so).
 *
 * Drop-in replacement for the device methods of Canary.jl's Euler (set2c)
 * and shallow-water examples.  All arrays use the reference (Julia,
 * column-major) layouts, all maps hold the reference's 1-based Int64 values,
 * and all device pointers are owned by the caller.
 *
 *   Q, rhs      (Np, nstate, nelem)          Q[n + Np*(s + nstate*e)]
 *   vgeo        (Np, nvgeo, nelem)           3D fields: xi_x eta_x zeta_x xi_y
 *                                            eta_y zeta_y xi_z eta_z zeta_z MJ
 *                                            MJI x y z (14); 2D: xi_x eta_x xi_y
 *                                            eta_y MJ MJI x y (8)
 *   sgeo        (nsgeo, Nfp, nface, nelem)   3D: nx ny nz sMJ vMJI (5);
 *                                            2D: nx ny sMJ vMJI (4)
 *   vmapM/P     (Nfp, nface, nelem) int64    1-based global node ids
 *   elemtobndy  (nface, nelem) int64         0 = connected, 1 = no-flux wall
 *
 * vmapM may be NULL in canary_facerhs, canary_stage_fused,
 * canary_facerhs_update and canary_lsrk54: the M-side ids are then the
 * canonical face mask, vmapM[:,f,e] = Np*(e-1) + fmask[:,f], the only vmapM
 * the reference's `mappings` builds (src/mesh.jl:973-1001, :994), and the
 * map is not read (one Int64 per face node less to move; same results).
 *   D           (Nq, Nq) column-major        D[j + Nq*k] = l_k'(r_j)
 *
 * Node n of an element is n = i + Nq*j (+ Nq^2*k), 0-based here.
 * Return values: 0 on success, a CANARY_ERR_* code otherwise.
 */
#ifndef CANARY_DG_H
#define CANARY_DG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes -------------------------------------------------------- */
#define CANARY_OK 0
#define CANARY_ERR_ARG 1            /* bad argument (size, pointer, N)        */
#define CANARY_ERR_UNSUPPORTED 2    /* (eqn, dim, N, prec) not compiled in    */
#define CANARY_ERR_LAUNCH 3         /* HIP launch / runtime error             */
#define CANARY_ERR_BC 4             /* elemtobndy value other than 0 or 1     */
#define CANARY_ERR_VMAP 5           /* vmapP outside [1, nelem_total*Np]      */
#define CANARY_ERR_NOCONVERGENCE 6  /* QL iteration did not converge          */
#define CANARY_ERR_LOBATTO_N1 7     /* Lobatto rule asked for n = 1           */
#define CANARY_ERR_ORDERING 8       /* face vertices with no known ordering   */
#define CANARY_ERR_ORIENTATION 9    /* 3D face orientation other than 1       */
#define CANARY_ERR_NOMEM 10         /* allocation failed                      */
#define CANARY_ERR_CAPACITY 11      /* request exceeds a stated fixed limit   */

/* ---- equation sets, precisions ----------------------------------------- */
#define CANARY_EULER 0   /* set2c: (rho u, rho v[, rho w], rho, rho theta)  */
#define CANARY_SWE 1     /* (H u, H v, h, b), b read only                    */
#define CANARY_EULER_TE 2 /* 2D total energy (set3c): (rho u, rho v, rho, E)  */
#define CANARY_FP64 0
#define CANARY_FP32 1
/* FP64 with a*b+c contracted into fused multiply-adds and 1/rho as a
 * Newton-refined reciprocal: fewer FP64 instructions, not the reference's
 * bit pattern (within 1e-12 relative L2 per stage of it; the Euler
 * pressure's pow as exp(kappa log x) by series, exact at x = 1).  Accepted
 * by every FP64 entry point (the three drop-in kernels, the fused ones and
 * the library loops); CANARY_FP64 is the exact reference order, the
 * pressure's pow correctly rounded. */
#define CANARY_FP64_FMA 2
/* CANARY_FP64_FMA with the volume term's derivative contractions (the sums
 * over D along xi, eta, zeta) on the matrix cores (FP64 MFMA, 16x16x4
 * tiles): 3D Euler only, accepted by canary_volumerhs, canary_stage_fused,
 * canary_stage_prepared and canary_lsrk54 / canary_lsrk54_prepared (whose
 * three-kernel and face + update paths run facerhs, updatesolution and that
 * fusion contracted: they have no derivative contraction; called directly
 * with this prec they report CANARY_ERR_UNSUPPORTED).  Within the contracted
 * form's tolerance of the reference.  A tile does useful work in at most half
 * of its multiply-adds (NB = 16 / (N+1) lines side by side), and on MI355X
 * the FP64 matrix rate equals the FP64 vector rate. */
#define CANARY_FP64_MFMA 3

/* Physical constants passed at run time (reference defaults in
 * canary_phys_default: examples/3d_kernels/euler3d_set2c.jl:58-63 and
 * examples/old_fxg_files/shallow_water.jl main()). */
typedef struct canary_phys {
  double gamma;    /* wave-speed gamma (reference: 14//10)               */
  double p0;       /* reference pressure (100000)                        */
  double R_gas;    /* 28717//100                                         */
  double c_p;      /* 100467//100                                        */
  double c_v;      /* 7175//10                                           */
  double gravity;  /* Euler buoyancy g (reference 10; bubble config 9.81) */
  double swe_gravity; /* SWE gravity (reference 10)                      */
  double delta_nl;    /* SWE nonlinear switch (reference 1)              */
} canary_phys;

void canary_phys_default(canary_phys* p);

/* ======================================================================
 * Drop-in device kernels (examples/3d_kernels/euler3d_set2c.jl,
 * examples/2d_kernels/euler2d_set2c.jl, examples/old_fxg_files/shallow_water.jl)
 * nreal:       number of real elements (reference `length(elems)`, elements
 *              1:nreal); ghost elements follow them in Q.
 * nelem_total: size(Q, 3) including ghosts; bounds every vmapP gather.
 * stream:      hipStream_t (NULL = default stream).  Launches are async.
 * Errors detected on the device (bad bc, vmapP out of range) are recorded
 * in a device-side status word and returned by canary_check_errors().
 * ====================================================================== */

/* volumerhs!(Val(dim), Val(N), rhs, Q, vgeo, D, elems[, gravity, δnl])
 *   replaces euler3d_set2c.jl:713-718 (CPU :144-231),
 *            euler2d_set2c.jl:598-603 (CPU :140-205),
 *            shallow_water.jl:560-563 (CPU :121-173),
 *            euler2d_set3c.jl:975-980 (CPU :214-279; eqn CANARY_EULER_TE).
 *   rhs += weak-form volume term (+ buoyancy for Euler). */
int canary_volumerhs(int eqn, int dim, int N, int prec, void* rhs, const void* Q,
                     const void* vgeo, const double* D, int64_t nreal,
                     const canary_phys* phys, void* stream);

/* facerhs!(Val(dim), Val(N), rhs, Q, sgeo, elems, vmapM, vmapP, elemtobndy)
 *   replaces euler3d_set2c.jl:720-726 (CPU :234-348),
 *            euler2d_set2c.jl:605-611 (CPU :208-300),
 *            shallow_water.jl fluxrhs! :565-569 (CPU :178-263).
 *   rhs -= sMJ * Rusanov flux, owner computes, no atomics. */
int canary_facerhs(int eqn, int dim, int N, int prec, void* rhs, const void* Q,
                   const void* sgeo, int64_t nreal, int64_t nelem_total,
                   const int64_t* vmapM, const int64_t* vmapP,
                   const int64_t* elemtobndy, const canary_phys* phys, void* stream);
/* fluxrhs!(Val(dim), Val(N), rhs, Q, sgeo, vgeo, elems, vmapM, vmapP,
 * elemtobndy) of the total-energy equations (eqn = CANARY_EULER_TE)
 *   replaces euler2d_set3c.jl:982-988 (CPU :282-378): the pressure reads
 *   vgeo's y at the M and P nodes, so vgeo (of all nelem_total elements) is
 *   required.  For the other equation sets vgeo may be NULL and the call is
 *   canary_facerhs. */
int canary_facerhs_geo(int eqn, int dim, int N, int prec, void* rhs, const void* Q,
                       const void* sgeo, const void* vgeo, int64_t nreal, int64_t nelem_total,
                       const int64_t* vmapM, const int64_t* vmapP, const int64_t* elemtobndy,
                       const canary_phys* phys, void* stream);

/* updatesolution!(Val(dim), Val(N), rhs, Q, vgeo, elems, rka, rkb, dt[, advection])
 *   replaces euler3d_set2c.jl:728-734 (CPU :352-358),
 *            euler2d_set2c.jl:613-619 (CPU :304-310),
 *            shallow_water.jl:571-575 (CPU :267-298).
 *   Q += rkb*dt*rhs*MJI; rhs *= rka.  `advection` (SWE only) freezes the
 *   velocity as the SWE CPU method does. */
int canary_updatesolution(int eqn, int dim, int N, int prec, void* rhs, void* Q,
                          const void* vgeo, int64_t nreal, double rka, double rkb,
                          double dt, int advection, void* stream);

/* One LSRK stage in a single pass: volume + faces + update, fused.
 *   Equivalent to volumerhs! ; facerhs! ; updatesolution! of one iteration of
 *   the stage loop (euler3d_set2c.jl:840-854), except that Q is read from
 *   Qin and the updated state is written to Qout (Qin != Qout; the caller
 *   swaps them each stage), so neighbouring elements never read a Q that
 *   is being overwritten.  Ghost elements (nreal..nelem_total-1) of Qin are
 *   read, never written.  rhs is updated in place.
 *   prec: CANARY_FP64 (bit-identical to the three calls above) or
 *   CANARY_FP64_FMA (contracted: fewer FP64 instructions; on the 3D N=4
 *   benchmark both run at the same speed, the stage being bound by HBM).
 *   advection: shallow water only, as in canary_updatesolution (the
 *   velocity kept through the update, shallow_water.jl:274-297). */
int canary_stage_fused(int eqn, int dim, int N, int prec, void* rhs, const void* Qin,
                       void* Qout, const void* vgeo, const void* sgeo,
                       const double* D, int64_t nreal, int64_t nelem_total,
                       const int64_t* vmapM, const int64_t* vmapP,
                       const int64_t* elemtobndy, double rka, double rkb, double dt,
                       int advection, const canary_phys* phys, void* stream);

/* facerhs! ; updatesolution! of one stage in a single pass, after a
 * separate canary_volumerhs (the opt-in face + update fusion):
 *   equivalent to the facerhs! and updatesolution! calls of one iteration of
 *   the stage loop (euler3d_set2c.jl:849-854), with Q ping-pong as in
 *   canary_stage_fused (Qin != Qout; ghosts of Qin read, never written).
 *   prec: CANARY_FP64 (bit-identical to canary_facerhs + canary_updatesolution)
 *   or CANARY_FP64_FMA (contracted).  advection as in canary_stage_fused. */
int canary_facerhs_update(int eqn, int dim, int N, int prec, void* rhs, const void* Qin,
                          void* Qout, const void* vgeo, const void* sgeo, int64_t nreal,
                          int64_t nelem_total, const int64_t* vmapM, const int64_t* vmapP,
                          const int64_t* elemtobndy, double rka, double rkb, double dt,
                          int advection, const canary_phys* phys, void* stream);

/* ---- prepared geometry (opt-in) -------------------------------------------
 * Compact copies of the reference arrays, built once per mesh, that the
 * fused stage reads instead of vgeo / sgeo / vmapM / vmapP / elemtobndy:
 * 34,824 bytes per element-stage at 3D N=4 against 38,248 with the
 * reference arrays.  No reference counterpart (the reference kernels read
 * the arrays computegeometry and mappings build, euler3d_set2c.jl:101-139,
 * src/mesh.jl:973-1001); everything left out is checked to be unused by
 * the stage or a function of what is kept:
 *   pvgeo double (Np, dim*dim+1, nelem): metric terms and MJ (MJI = 1 ./ MJ
 *         checked bit for bit and recomputed by the stage; x, y, z unused)
 *   psgeo double (dim+1, Nfp, 2 dim, nelem): normal and sMJ (vMJI unused)
 *   pface int32 (2 dim, nelem): -1 for a wall (elemtobndy 1), else
 *         2 dim * eP + fP: vmapP[:, f, e] = Np eP + fmask[:, fP] + 1 checked
 *         for every node (the only orientation mappings builds, mesh.jl:991)
 * vmapM must be the canonical face mask and elemtobndy 0 or 1.  A failed
 * check is reported by canary_check_errors (CANARY_ERR_ARG for MJI,
 * CANARY_ERR_VMAP, CANARY_ERR_ORIENTATION, CANARY_ERR_BC); the outputs are
 * then not to be used.  Ghost elements (nreal..nelem_total-1) may be
 * neighbours. */
int canary_prepare_geometry(int dim, int N, int64_t nelem, int64_t nelem_total, const void* vgeo,
                            const void* sgeo, const int64_t* vmapM, const int64_t* vmapP,
                            const int64_t* elemtobndy, void* pvgeo, void* psgeo, int32_t* pface,
                            void* stream);
/* canary_stage_fused over the prepared arrays: the same arithmetic in the
 * same order (prec CANARY_FP64: bit-identical to the three reference calls;
 * CANARY_FP64_FMA: the contracted form), fewer bytes. */
int canary_stage_prepared(int eqn, int dim, int N, int prec, void* rhs, const void* Qin,
                          void* Qout, const void* pvgeo, const void* psgeo, const int32_t* pface,
                          const double* D, int64_t nreal, int64_t nelem_total, double rka,
                          double rkb, double dt, int advection, const canary_phys* phys,
                          void* stream);

/* ---- native time loop ---------------------------------------------------
 * lowstorageRK (euler3d_set2c.jl:754-891, euler2d_set2c.jl:639-775,
 * shallow_water.jl:615-747): nsteps LSRK54 steps of one solve through the
 * entry points above, stage loop as the reference's (:819-855).
 *   path CANARY_PATH_KERNELS: volumerhs, facerhs, updatesolution per stage,
 *        Q updated in place (Qalt unused), prec CANARY_FP64 only;
 *   path CANARY_PATH_FUSED: canary_stage_fused, Q and Qalt ping-pong;
 *   path CANARY_PATH_FACE_UPDATE: canary_volumerhs + canary_facerhs_update,
 *        Q and Qalt ping-pong.
 *   `advection` (shallow water, any path) as in canary_updatesolution.
 *   The loop exchanges no ghost elements, so nelem_total must equal nreal
 *   (CANARY_ERR_ARG otherwise); a sharded caller runs its own stage loop.
 *   *result_in_alt: 1 if the final state is in Qalt, else 0 (in Q).
 *   graph != 0 and a non-null stream: two steps are captured once into a
 *   hipGraph and replayed (the call then returns after the steps ran). */
#define CANARY_PATH_KERNELS 0
#define CANARY_PATH_FUSED 1
#define CANARY_PATH_FACE_UPDATE 2
int canary_lsrk54(int eqn, int dim, int N, int prec, int path, int64_t nsteps, double dt,
                  void* Q, void* Qalt, void* rhs, const void* vgeo, const void* sgeo,
                  const double* D, int64_t nreal, int64_t nelem_total, const int64_t* vmapM,
                  const int64_t* vmapP, const int64_t* elemtobndy, const canary_phys* phys,
                  int advection, int graph, int* result_in_alt, void* stream);
/* canary_lsrk54 over canary_stage_prepared (Q and Qalt ping-pong; no ghost
 * elements, as canary_lsrk54). */
int canary_lsrk54_prepared(int eqn, int dim, int N, int prec, int64_t nsteps, double dt, void* Q,
                           void* Qalt, void* rhs, const void* pvgeo, const void* psgeo,
                           const int32_t* pface, const double* D, int64_t nreal,
                           const canary_phys* phys, int advection, int graph, int* result_in_alt,
                           void* stream);
/* Makes the table of the correctly rounded pow for one gas exponent
 * kappa = cp/cv on the current device (once; kappa in [1, 2], up to 16 of
 * them).  The exact-order Euler entries make it themselves on first use;
 * a caller that captures their launches into a graph, or runs batches
 * (whose constants live in device memory), registers each kappa first.
 * Without a table the kernels give the same bits more slowly.  0, or
 * CANARY_ERR_UNSUPPORTED when no table was made.  No reference counterpart. */
int canary_pow_register(double kappa, void* stream);
/* RKA, RKB, RKC of LSRK54 as the reference's rationals (euler3d_set2c.jl:772-788);
 * any pointer may be NULL. */
void canary_lsrk54_coefs(double* rka, double* rkb, double* rkc);

/* ---- precision form (Euler set2c, dim 2 or 3, N = 1..7) -------------------
 * FP32 perturbation q' = Q - Q0 about an FP64 background Q0 at rest
 * (hydrostatic, or uniform without gravity); no reference counterpart: the
 * reference evolves Q in one precision (euler3d_set2c.jl:754-891,
 * euler2d_set2c.jl:639-775).  Arrays (column-major, per element; NS = dim+2,
 * NF = 2 dim):
 *   vg32 float (Np,dim*dim+1,nelem) metric, MJ       sg32 float (4,Nfp,NF,nelem) nx ny nz sMJ
 *   nbr  int32 (NF,nelem) face code: -1 wall, else NF*eP + fP (0-based
 *        neighbour element eP and face fP)
 *   bg   double (Np,3,nelem) rho0, rhotheta0, P0       bg32 float (Np,3,nelem) the same in FP32
 *   qp, rhs float (Np,NS,nelem)
 * prepare: builds them from the FP64 drop-in arrays, the background Q0 and a
 * full state Q (qp = Q - Q0).  It checks what the compact arrays leave out
 * (vmapM the canonical face mask, each vmapP column the canonical mask of
 * one neighbour face, elemtobndy 0 or 1, MJI == 1 ./ MJ) and records the
 * first violation (CANARY_ERR_VMAP, _ORIENTATION, _BC, _ARG) for
 * canary_check_errors.  The stage reads bg32 (the FP32 values its
 * arithmetic uses); reconstruct reads bg. */
int canary_perturb32_prepare(int dim, int N, int64_t nelem, int64_t nelem_total,
                             const void* vgeo, const void* sgeo, const int64_t* vmapM,
                             const int64_t* vmapP, const int64_t* elemtobndy, const void* Q0,
                             const void* Q, const canary_phys* phys, void* vg32, void* sg32,
                             int32_t* nbr, void* bg, void* bg32, void* qp, void* stream);
/* one fused LSRK stage of the perturbation (qin -> qout, rhs in place);
 * MJI = 1 / MJ in the FP64 update */
int canary_stage_perturb32(int dim, int N, void* rhs, const void* qin, void* qout,
                           const void* bg32, const void* vg32, const void* sg32,
                           const int32_t* nbr, const double* D, int64_t nreal, double rka,
                           double rkb, double dt, const canary_phys* phys, void* stream);
/* Q (FP64, (Np,NS,nelem)) = Q0 + qp (both variants; bg the FP64 background) */
int canary_perturb32_reconstruct(int dim, int N, int64_t nelem, const void* qp, const void* bg,
                                 void* Q, void* stream);
/* Mixed variant: the same FP32 state q' and FP32 RHS arithmetic, with FP64
 * kept where the compact arrays round: metric terms and MJ
 *   vg64 double (Np,dim*dim+1,nelem),  sg64 double (4,Nfp,NF,nelem) nx ny nz sMJ,
 * and the LSRK register rhs64 double (Np,NS,nelem) (dq = dq + L(q') and the
 * update q' + ((rkb*dt) dq) MJI in FP64 with MJI = 1 / MJ, the reference's
 * value, q' rounded to FP32 once).  No reference counterpart (see the
 * precision form above). */
int canary_perturb_mixed_prepare(int dim, int N, int64_t nelem, int64_t nelem_total,
                                 const void* vgeo, const void* sgeo, const int64_t* vmapM,
                                 const int64_t* vmapP, const int64_t* elemtobndy, const void* Q0,
                                 const void* Q, const canary_phys* phys, void* vg64, void* sg64,
                                 int32_t* nbr, void* bg, void* bg32, void* qp, void* stream);
int canary_stage_perturb_mixed(int dim, int N, void* rhs64, const void* qin, void* qout,
                               const void* bg32, const void* vg64, const void* sg64,
                               const int32_t* nbr, const double* D, int64_t nreal, double rka,
                               double rkb, double dt, const canary_phys* phys, void* stream);

/* ---- batched form ---------------------------------------------------------
 * Many independent solves of one equation set / dimension, each with its own
 * N, element count, mesh, dt and constants, advanced through a device work
 * table of element blocks.  No reference counterpart (the reference runs one
 * solve per process, euler3d_set2c.jl:1049).  Arrays of a solve are the
 * drop-in ones (FP64: prec = CANARY_FP64 in the reference evaluation order
 * or CANARY_FP64_FMA contracted, as for canary_stage_fused), or FP32 copies
 * of Q, rhs, vgeo and sgeo for prec = CANARY_FP32 (shallow water only).
 * A solve in the prepared layout (canary_prepare_geometry's arrays) has
 * pface set, pvgeo / psgeo in vgeo / sgeo (FP32 copies for CANARY_FP32),
 * and vmapM, vmapP, elemtobndy unused; it runs through the *_prepared
 * entries.  Both tables live in device memory. */
typedef struct canary_batch_solve {
  void* Q[2];                  /* ping-pong states; stage reads Q[parity], writes Q[1-parity] */
  void* rhs;                   /* LSRK register                                  */
  const void* vgeo;
  const void* sgeo;
  const double* D;
  const int64_t* vmapM;
  const int64_t* vmapP;
  const int64_t* elemtobndy;
  int64_t nreal, nelem_total;
  double dt;
  canary_phys phys;
  int32_t N;
  int32_t advection;           /* shallow water: keep the velocity (as updatesolution) */
  const int32_t* pface;        /* prepared layout: face codes; NULL: reference arrays */
} canary_batch_solve;

typedef struct canary_batch_item {
  int32_t solve;  /* index into the solve table          */
  int16_t N;      /* polynomial order of that solve      */
  int16_t epw;    /* elements per workgroup (variant)    */
  int64_t e0;     /* first element of the block (0-based) */
} canary_batch_item;

/* Candidate `which` (0, 1, 2) of elements-per-workgroup for (eqn, dim, prec,
 * N): writes epw and the threads per workgroup of that variant; returns
 * CANARY_ERR_UNSUPPORTED for a candidate that does not exist. */
int canary_batch_variant(int eqn, int dim, int prec, int N, int which, int* epw, int* threads);
/* One LSRK stage of the `nitems` work-table entries at `items`, all of
 * order N and variant epw: Q[parity] -> Q[1-parity] of their solves.  An
 * entry of another N / epw, or an e0 outside its solve, is skipped and
 * recorded as CANARY_ERR_ARG (canary_check_errors).  One workgroup per
 * entry; consecutive entries land on the same XCD (shared L2). */
int canary_batch_stage(int eqn, int dim, int prec, int N, int epw,
                       const canary_batch_solve* solves, const canary_batch_item* items,
                       int64_t nitems, int parity, double rka, double rkb, void* stream);

/* The batched stage in ONE launch for entries of every order: each
 * workgroup (256 lanes) takes one work-table entry and runs the stage body
 * of that entry's N.  Every entry's epw must be the one
 * canary_batch_all_variant gives for its N (the most elements that fit 256
 * lanes); an entry of another epw, of an order that does not fit (3D
 * N = 6, 7) or outside its solve is skipped and recorded as CANARY_ERR_ARG.
 * The kernel is sized by its largest order (registers and LDS), so it runs
 * fewer waves per CU than the per-order kernels of canary_batch_stage; it
 * takes the stage in one launch instead of one per order.  No reference
 * counterpart (the reference runs one solve per process). */
int canary_batch_all_variant(int eqn, int dim, int prec, int N, int* epw, int* threads);
int canary_batch_stage_all(int eqn, int dim, int prec, const canary_batch_solve* solves,
                           const canary_batch_item* items, int64_t nitems, int parity,
                           double rka, double rkb, void* stream);

/* The same two launch forms over solves in the prepared layout (every
 * solve of the table with pface set; an entry of a solve without it is
 * skipped and recorded as CANARY_ERR_ARG, as is a prepared solve's entry
 * in canary_batch_stage / canary_batch_stage_all). */
int canary_batch_stage_prepared(int eqn, int dim, int prec, int N, int epw,
                                const canary_batch_solve* solves, const canary_batch_item* items,
                                int64_t nitems, int parity, double rka, double rkb, void* stream);
int canary_batch_stage_all_prepared(int eqn, int dim, int prec, const canary_batch_solve* solves,
                                    const canary_batch_item* items, int64_t nitems, int parity,
                                    double rka, double rkb, void* stream);

/* Batched precision form: one solve of the table is the set of arrays of
 * the precision form above (q[parity] -> q[1-parity]); items as for
 * canary_batch_stage. */
typedef struct canary_p32_solve {
  void* q[2];
  void* rhs;
  const void* bg32;
  const void* vg;
  const void* sg;
  const int32_t* nbr;
  const double* D;
  int64_t nreal;
  double dt;
  canary_phys phys;
  int32_t N;
  int32_t dim;
} canary_p32_solve;

int canary_p32_batch_variant(int dim, int N, int which, int* epw, int* threads);
int canary_p32_batch_stage(int dim, int N, int epw, const canary_p32_solve* solves,
                           const canary_batch_item* items, int64_t nitems, int parity,
                           double rka, double rkb, void* stream);

/* nsteps LSRK54 steps of a whole batch from C: for every stage (the
 * reference's stage loop, euler3d_set2c.jl:819-855), each class's entries in
 * turn, then the parity flips.  kind CANARY_BATCH_DROPIN: canary_batch_stage
 * with (eqn, dim, prec) and a canary_batch_solve table; kind
 * CANARY_BATCH_P32: canary_p32_batch_stage with dim and a canary_p32_solve
 * table.  `classes` is a host array; its items pointers are device memory.
 * No reference counterpart: it takes the per-stage host loop off the
 * launch path for batches of many small solves.  Every class is checked
 * (compiled (N, epw) variant, items pointer) before the first launch, so an
 * argument error launches nothing; a launch error returns at once. */
typedef struct canary_batch_class {
  const canary_batch_item* items; /* this class's work-table entries (device) */
  int64_t nitems;
  int32_t N;
  int32_t epw;
} canary_batch_class;
#define CANARY_BATCH_DROPIN 0
#define CANARY_BATCH_P32 1
#define CANARY_BATCH_DROPIN_ALL 2 /* canary_batch_stage_all per class (N, epw ignored) */
#define CANARY_BATCH_PREPARED 3   /* canary_batch_stage_prepared */
#define CANARY_BATCH_PREPARED_ALL 4 /* canary_batch_stage_all_prepared (N, epw ignored) */
int canary_batch_run(int kind, int eqn, int dim, int prec, const void* solves,
                     const canary_batch_class* classes, int nclasses, int64_t nsteps, int parity,
                     void* stream);

/* Synchronises `stream`, returns the first recorded device error
 * (CANARY_ERR_BC / CANARY_ERR_VMAP or 0), writes (code, face, element,
 * value) of it to info[0..3] if info != NULL, and clears the record. */
int canary_check_errors(void* stream, int64_t* info);

/* Number of HIP devices visible (0 without a GPU); never faults. */
int canary_device_count(void);

/* A stream of the caller's own (non-blocking, default priority) and its
 * release (after the work on it completes); used by the batched and served
 * runtimes so that their streams never alias one another. */
int canary_stream_create(void** stream);
int canary_stream_destroy(void* stream);

/* ======================================================================
 * Host preparation (the inputs of the kernels).  All outputs are written
 * to caller-provided host arrays in the reference layouts.
 * ====================================================================== */

#define CANARY_ENDPT_NEITHER 0
#define CANARY_ENDPT_LEFT 1
#define CANARY_ENDPT_RIGHT 2
#define CANARY_ENDPT_BOTH 3

/* GaussQuadrature.legendre(T, n, endpt) with maxits (<=0: 30):
 * src/GaussQuadrature/quadrature.jl:126-214 */
int canary_gauss_rule(int n, int endpt, int maxits, double* x, double* w);
/* lglpoints / lgpoints: src/operators.jl:10-25 */
int canary_lglpoints(int N, double* r, double* w);
int canary_lgpoints(int N, double* r, double* w);
/* baryweights / spectralderivative / interpolationmatrix:
 * src/operators.jl:56-69, 84-104, 118-137 */
int canary_baryweights(int n, const double* r, double* wb);
int canary_spectralderivative(int n, const double* r, double* D);
int canary_interpolationmatrix(int nsrc, const double* rsrc, int ndst,
                               const double* rdst, double* I);

/* hilbertcode(Y; bits): src/mesh.jl:35-88 (X, H are n words) */
int canary_hilbertcode(int n, const uint64_t* Y, int bits, uint64_t* H);
/* centroidtocode(comm, elemtocoord) on one rank: src/mesh.jl:107-141.
 * elemtocoord (d, nvert, nelem) -> code (d, nelem) */
int canary_centroidtocode(int d, int nvert, int64_t nelem, const double* elemtocoord,
                          uint64_t* code);
/* vertsortandorder(v...) for n = 1..4 vertices: src/mesh.jl:595-719 */
int canary_vertsortandorder(int n, const int64_t* v, int64_t* sorted, int* order);

/* Opaque mesh built by brickmesh -> partition -> connectmesh on one rank. */
typedef struct canary_mesh canary_mesh;

/* brickmesh(x, periodic; part, numparts, boundary): src/mesh.jl:259-317.
 * ncoord[j] = length(x[j]); coords = x[1], x[2], ... concatenated;
 * boundary = (2, d) column-major or NULL for all ones. */
canary_mesh* canary_brickmesh(int d, const int64_t* ncoord, const double* coords,
                              const int* periodic, int64_t part, int64_t numparts,
                              const int64_t* boundary, int* status);
/* A mesh from the four arrays brickmesh returns (faceconnections as
 * (2+2^(d-1), nfc) column-major), e.g. to partition a user-built mesh. */
canary_mesh* canary_mesh_from_arrays(int d, int64_t nelem, const int64_t* elemtovert,
                                     const double* elemtocoord, const int64_t* elemtobndy,
                                     int64_t nfc, const int64_t* faceconnections, int* status);
/* partition(comm, mesh...) on one rank (Hilbert order): src/mesh.jl:499-580 */
int canary_mesh_partition(canary_mesh* m);
/* connectmesh(comm, mesh...) on one rank: src/mesh.jl:751-956 */
int canary_mesh_connect(canary_mesh* m);
/* sizes: what = 0 nelem, 1 dim, 2 nfaceconnections, 3 connected(0/1) */
int64_t canary_mesh_size(const canary_mesh* m, int what);
/* copy an array out: name is one of "elemtovert" (nvert,nelem),
 * "elemtocoord" (d,nvert,nelem, double), "elemtobndy" (nface,nelem),
 * "elemtoelem", "elemtoface", "elemtoordr" (nface,nelem),
 * "faceconnections" (2+2^(d-1), nfc) -- int64 unless stated. */
int canary_mesh_get(const canary_mesh* m, const char* name, void* out);
void canary_mesh_free(canary_mesh* m);

/* mappings(N, elemtoelem, elemtoface, elemtoordr): src/mesh.jl:973-1001 */
int canary_mappings(int N, int d, int64_t nelem, const int64_t* elemtoelem,
                    const int64_t* elemtoface, const int64_t* elemtoordr,
                    int64_t* vmapM, int64_t* vmapP);

/* creategrid!(x[, y[, z]], elemtocoord, r): src/metric.jl:11-107.
 * X = d arrays of Nq^d*nelem doubles, contiguous one after the other. */
int canary_creategrid(int d, int Nq, int64_t nelem, const double* elemtocoord,
                      const double* r, double* X);
/* computemetric!(...): src/metric.jl:120-400.
 * X: d coordinate arrays; vol: J then the d*d metric terms in the
 * reference argument order (1D: xi_x; 2D: xi_x eta_x xi_y eta_y;
 * 3D: xi_x eta_x zeta_x xi_y eta_y zeta_y xi_z eta_z zeta_z), each
 * Nq^d*nelem; surf: sJ then n_x .. n_d, each Nq^(d-1)*2d*nelem. */
int canary_computemetric(int d, int Nq, int64_t nelem, const double* X, const double* D,
                         double* vol, double* surf);
/* computegeometry(Val(dim), mesh, D, ξ, ω, identity, vmapM):
 * euler3d_set2c.jl:101-139 / euler2d_set2c.jl:97-134, no mesh warp.
 * Writes vgeo (Np, nvgeo, nelem) and sgeo (nsgeo, Nfp, nface, nelem). */
int canary_computegeometry(int d, int N, int64_t nelem, const double* elemtocoord,
                           const int64_t* vmapM, double* vgeo, double* sgeo);

/* ---- initial conditions, dt, energy (host) ------------------------------ */

/* Rising thermal bubble (euler3d_set2c.jl:1189-1218, euler2d_set2c.jl:1056-1083):
 * theta = theta0 + 0.5*dtheta*(1+cos(pi r/rc)) for r <= rc (+ jitter). */
typedef struct canary_bubble {
  double theta0;     /* background potential temperature (300)            */
  double dtheta;     /* theta_c: peak perturbation (reference 0.5)        */
  double rc;         /* radius (250)                                      */
  double xc, yc, zc; /* centre; zc is the vertical centre (350)           */
  int spherical;     /* 3D only: 1 = sphere, 0 = the reference x-z cylinder */
  double jitter;     /* amplitude of a seeded theta jitter (0 = none)     */
  uint64_t seed;
} canary_bubble;

/* Isentropic vortex (no reference counterpart; exact solution). */
typedef struct canary_vortex {
  double beta, x0, y0, u0, v0, rho_inf, p_inf, gamma, Lx, Ly;
} canary_vortex;

/* x^y correctly rounded to double (positive finite x, finite y,
 * |y log x| < 700; outside that, the C library's pow): the pow the
 * exact-order kernels use for the pressure, here on the host.  The
 * reference calls Julia's libm pow (euler3d_set2c.jl:175), which no test
 * fixes to the last bit; this is one function of (x, y) on host and device. */
double canary_pow_cr(double x, double y);
/* canary_pow_cr through the kernels' fixed-exponent table of y (built on
 * the host for y in [1, 2]); the same value, to check that path. */
double canary_pow_cr_table(double x, double y);
double canary_hash_jitter(uint64_t seed, double x, double y, double z);
int canary_ic_bubble(int dim, int N, int64_t nelem, const double* vgeo,
                     const canary_bubble* b, const canary_phys* ph, double* Q);
int canary_ic_vortex2d(int N, int64_t nelem, const double* vgeo, const canary_vortex* v,
                       double t, double* Q);
/* bumps: nbump records (xc, yc, amplitude, width) */
int canary_ic_swe_bumps(int N, int64_t nelem, const double* vgeo, double depth, int nbump,
                        const double* bumps, double Lx, double Ly, double* Q);
/* cfl(...) before the `/ N^sqrt(2)` (euler3d_set2c.jl:67-97, euler2d_set2c.jl:66-93,
 * shallow_water.jl:43-67) */
double canary_cfl(int eqn, int dim, int N, const double* vgeo, const double* Q,
                  int64_t nelem, const canary_phys* ph);
/* L2energysquared (euler3d_set2c.jl:738-750) */
double canary_l2energysquared(int Np, int nstate, int nvgeo, int iMJ, const double* Q,
                              const double* vgeo, int64_t nelem);

/* ---- device side of the driver's other work (device pointers) ------------ */

/* computegeometry on the GPU (euler3d_set2c.jl:101-139 / euler2d_set2c.jl:97-134):
 * elemtocoord (d, 2^d, nelem), vmapM (1-based) -> vgeo, sgeo, bit-identical
 * to canary_computegeometry.  Errors (vmapM out of range) are recorded for
 * canary_check_errors. */
int canary_computegeometry_device(int d, int N, int64_t nelem, const double* elemtocoord,
                                  const int64_t* vmapM, double* vgeo, double* sgeo,
                                  void* stream);

/* Ghost-element exchange buffers (word_bytes 8 = FP64 states, 4 = FP32):
 * fillsendQ!:     sendQ[:, :, e] = Q[:, :, sendelems[e]], sendelems 1-based
 *                 (euler3d_set2c.jl:645-659, 682-693);
 * transferrecvQ!: Q[:, :, nrealelem + e] = recvQ[:, :, e]
 *                 (euler3d_set2c.jl:663-676, 695-709). */
int canary_fillsendQ(int dim, int N, int nstate, void* sendQ, const void* Q,
                     const int64_t* sendelems, int64_t nsend, int64_t nelem_total,
                     int word_bytes, void* stream);
int canary_transferrecvQ(int dim, int N, int nstate, void* Q, const void* recvQ, int64_t nrecv,
                         int64_t nrealelem, int word_bytes, void* stream);

/* Face-trace halo: the nodes a neighbour reads through vmapP instead of
 * whole ghost elements (Nfp of Np nodes per shared face; no reference
 * counterpart, the reduced exchange of the sharded design).  nodes: 1-based
 * node ids over the (Np, nstate, nelem_total) array, like vmapP;
 * buf[k + nnode * s] = Q[node k, state s].  A node id out of range is
 * skipped and recorded as CANARY_ERR_VMAP. */
int canary_pack_nodes(int dim, int N, int nstate, void* buf, const void* Q, const int64_t* nodes,
                      int64_t nnode, int64_t nelem_total, int word_bytes, void* stream);
int canary_unpack_nodes(int dim, int N, int nstate, void* Q, const void* buf,
                        const int64_t* nodes, int64_t nnode, int64_t nelem_total,
                        int word_bytes, void* stream);

/* Deterministic device reductions into out[0] (device memory); work holds
 * CANARY_REDUCE_WORK doubles of scratch.  Same values as the host
 * canary_l2energysquared / canary_cfl up to summation order (energy) and
 * the last bit of pow (cfl, Euler). */
#define CANARY_REDUCE_WORK 1024
int canary_l2energysquared_device(int Np, int nstate, int nvgeo, int iMJ, const double* Q,
                                  const double* vgeo, int64_t nelem, double* work, double* out,
                                  void* stream);
int canary_cfl_device(int eqn, int dim, int N, const double* vgeo, const double* Q,
                      int64_t nelem, const canary_phys* phys, double* work, double* out,
                      void* stream);
/* canary_cfl_device with no `work` argument and no allocation: the
 * CANARY_REDUCE_WORK partials go to a static device array of the library
 * (8 KiB per device, held until the process ends; its address is looked up
 * on the first call made with each device current).  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * hipGraph capture, serving loops that admit requests without hipMalloc).
 * The partial count is fixed, so every nelem fits.  Calls that overlap on
 * one device share the array and race: issue them on one stream. */
int canary_cfl_device_noalloc(int eqn, int dim, int N, const double* vgeo, const double* Q,
                              int64_t nelem, const canary_phys* phys, double* out,
                              void* stream);
/* Courant-number step of the Euler sets: min over nodes of
 * courant * dx / (|u| + c) / N with dx = sqrt(sum_d (1/(2 xi^d_d))^2), the
 * form of the shallow-water example's courantnumber (swe2d.jl:100-124) with
 * the Euler wave speed; no reference counterpart for Euler. */
int canary_courant_device(int dim, int N, const double* vgeo, const double* Q, int64_t nelem,
                          const canary_phys* phys, double courant, double* work, double* out,
                          void* stream);
/* The host initial states of canary_ic_bubble / canary_ic_vortex2d evaluated
 * on the device from the coordinates in a device vgeo (same formulas; the
 * device cos / pow / exp may differ from the host's in the last bit). */
int canary_ic_bubble_device(int dim, int N, int64_t nelem, const double* vgeo,
                            const canary_bubble* b, const canary_phys* ph, double* Q,
                            void* stream);
int canary_ic_vortex2d_device(int N, int64_t nelem, const double* vgeo, const canary_vortex* v,
                              double t, double* Q, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* CANARY_DG_H */
