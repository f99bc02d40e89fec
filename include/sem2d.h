/*
 * sem2d.h -- C ABI of the MI355X explicit spectral-element time step.
 *
 * This is the drop-in boundary for the time loop of SEM2DPACK (SRC/main.f90:51-99):
 * the host (the reference's own Fortran initialisation, through ISO_C_BINDING, or
 * any other host) builds the problem once, hands its arrays to sem2d_create and
 * the sem2d_set_* calls, then advances the solution with sem2d_step and reads
 * results back with the sem2d_fetch_* calls.
 *
 * Conventions (those of the reference's Fortran arrays, so Fortran can pass its
 * arrays unchanged):
 *   - node indices are 1-based global GLL node numbers (SRC/spec_grid.f90:198-314);
 *   - nodal fields are [ndof][npoin] (Fortran field(npoin,ndof), column-major),
 *     so each component is contiguous;
 *   - ibool is [nelem][ngll][ngll] (Fortran ibool(i,j,e), i fastest);
 *   - element coefficients are [nelem][nelast][ngll][ngll] (Fortran a(i,j,k) of
 *     each element, SRC/mat_elastic.f90:11-14), a single [nelast][ngll][ngll]
 *     set shared by all elements (SRC/mat_gen.f90:357-365), or, on flat grids,
 *     the moduli per GLL node with three geometric factors per element;
 *   - hprime is [ngll][ngll] column-major: hprime[j*ngll+i] = h'_i(x_j).
 * Every call returns 0 on success and a non-zero code otherwise; the message is
 * available from sem2d_last_error (the reference prints and stops instead,
 * SRC/stdio.f90:205-214).  Host arrays are only read during the call.  One
 * handle owns one HIP stream and its device memory; a handle is not re-entrant.
 * ngll is 3 to 10.  At 5 (OPT_NGLL, SRC/constants.f90:6) the elements follow
 * ELAST_KD2_* and structured boxes use the fused tile kernels; at other values the
 * elements follow ELAST_KD1_* (SRC/mat_elastic.f90:464-518,645-681) on the general
 * element kernel (coef_mode 0 or 1; the compact mode 2 needs ngll = 5).  The batch and
 * the adjoint need ngll = 5.
 */
#ifndef SEM2D_H
#define SEM2D_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sem2d_solver sem2d_solver;

/* Problem description for sem2d_create. */
typedef struct sem2d_desc {
  int ndof;              /* 1 = SH (antiplane), 2 = P-SV (in-plane) */
  int ngll;              /* 3 .. 10 */
  int nelem, npoin;
  int nelast;            /* 2, 3 (SH flat/general), 6, 10 (P-SV flat/general) */
  int coef_mode;         /* 0: coefs per element; 1: one coefs set shared by all elements;
                            2: compact flat-grid form (lam_mu, geom, wgll; nelast 2 or 6) */
  const int *ibool;      /* [nelem][ngll][ngll], 1-based */
  const double *coefs;   /* modes 0, 1: [nelem or 1][nelast][ngll][ngll] */
  const double *lam_mu;  /* mode 2: [nelem][ndof][ngll][ngll]: lambda, mu (P-SV) or mu (SH) */
  const double *geom;    /* mode 2: [nelem][3]: -detJ*DxiDx^2, -detJ*DetaDz^2, -detJ*DxiDx*DetaDz,
                            so that a(i,j,k) = wgll(i)*wgll(j) * modulus(i,j) * factor
                            (SRC/mat_elastic.f90:334-340, 356-358) */
  const double *wgll;    /* mode 2: [ngll] GLL weights */
  const double *hprime;  /* [ngll][ngll], hprime[j*ngll+i] = h'_i(x_j) */
  const double *rmass;   /* [ndof][npoin], inverse of the assembled diagonal mass */
  double dt, beta, gamma;/* Newmark time step and coefficients (SRC/time.f90:5-11) */
  int bulk_fp32;         /* 0: FP64 fields and element update; 1: FP32 bulk update */
  /* Optional structured-grid description (a MESH_CART box); nx = 0 when absent.
     elem_of_cell[iz*nx+ix] is the 1-based element index of cell (ix, iz); ezflt is
     the element row under a split fault (0 for none), as in SRC/mesh_cartesian.f90. */
  int nx, nz, ezflt;
  const int *elem_of_cell;
} sem2d_desc;

/* Dynamic fault description (replaces the state of bc_dynflt_type, SRC/bc_dynflt.f90:17-38,
   as built by BC_DYNFLT_init, SRC/bc_dynflt.f90:231-520).  Arrays are [ncomp][npoin]. */
typedef struct sem2d_fault_desc {
  int npoin;                          /* fault node pairs */
  const int *node1, *node2;           /* 1-based nodes of side 1 and side 2; node2 = NULL
                                         (and invM2 = NULL) for a fault with one side, on a
                                         domain boundary (symmetry, SRC/bc_dynflt.f90:256-259,
                                         get_jump/get_weighted_jump :692-719) */
  const double *n1;                   /* [2][npoin] normal of side 1 */
  const double *B, *invM1, *invM2, *Z;/* [ndof][npoin] */
  const double *T0;                   /* [2][npoin] initial traction (tangent, normal) */
  const double *cohesion;             /* [npoin] */
  double coefA2V, coefA2D;
  int allow_opening;
  int law;                            /* 1 = slip weakening, 2 = rate and state, 3 = time
                                         weakening alone (sem2d_set_fault_twf) */
  int kind;                           /* SWF kind 1..3 (SRC/bc_dynflt_swf.f90:148-158),
                                         RSF kind 1..4 (SRC/bc_dynflt_rsf.f90:165-184) */
  int healing;                        /* SWF instantaneous healing */
  /* friction parameters, [npoin] each:
       SWF: p[0]=Dc p[1]=MuS p[2]=MuD p[3]=alpha p[4]=p
       RSF: p[0]=Dc p[1]=MuS p[2]=a   p[3]=b     p[4]=Vstar p[5]=Vc */
  const double *p[6];
  const double *theta0;               /* [npoin] initial state variable */
  const double *V0;                   /* [npoin] initial slip rate (RSF) or NULL */
  /* output decimation as in SRC/bc_dynflt.f90:419-433: records at steps
     it = oit, oit+oitd, ... over nodes oix1:oixn:oixd (1-based, inclusive) */
  int oit, oitd, oix1, oixn, oixd;
} sem2d_fault_desc;

/* Create a solver from the problem arrays: replaces pb%fields / compute_Fint set-up
   (SRC/fields.f90:52-90, SRC/mat_gen.f90:323-410).  Fields start at zero. */
int sem2d_create(const sem2d_desc *desc, sem2d_solver **out);

/* One absorbing boundary (Clayton-Engquist, flat), call order = BC application order:
   replaces BC_ABSO_apply (SRC/bc_abso.f90:286-336).  C is [ndof][n]. */
int sem2d_set_abso(sem2d_solver *h, int n, const int *node, const double *C);

/* Stacey term of the absorbing boundary registered last (P-SV, flat; BC_ABSORB
   stacey = T): replaces bc%K and the KxD part of BC_ABSO_apply (SRC/bc_abso.f90:176-224,
   321-333).  For boundary element e (boundary order) and edge point k (the order of the
   boundary's element edges), enode[e][k] is the 0-based index into that boundary's node
   array and K[e][k][c] the coefficient of component c; the force at the boundary nodes
   loses sum_e K[e][k][c] * sum_j h'_j(x_k) D_c(node[enode[e][j]]), D the predicted
   displacement, after the boundary's -C v term. */
int sem2d_set_abso_stacey(sem2d_solver *h, int nbe, const int *enode, const double *K);

/* The absorbing boundary registered last is not flat (its normals are not all along x
   or z; P-SV): replaces the non-flat branch of BC_ABSO_apply (SRC/bc_abso.f90:313-318).
   nrm is [2][n], the unit normal at each of its n nodes; its C is then [rho cp w|J|,
   rho cs w|J|] and the force loses C[0] (v.n) n + C[1] (v - (v.n) n).  Only for solvers
   on the general path (not a MESH_CART grid); n must be that boundary's node count. */
int sem2d_set_abso_normals(sem2d_solver *h, int n, const double *nrm);

/* Point force at a node: replaces FORCE_add (SRC/src_force.f90:77-90).  dir[2] is used
   when ndof = 2.  Returns the source index through *isrc. */
int sem2d_set_force(sem2d_solver *h, int iglob, const double *dir, int *isrc);

/* Moment tensor stencil: replaces SRC_MOMENT_add (SRC/src_moment.f90:217-235).
   iglob_xi/eta are [nel][ngll], coef_xi/eta are [nel][ndof][ngll]. */
int sem2d_set_moment(sem2d_solver *h, int nel, const int *iglob_xi, const int *iglob_eta,
                     const double *coef_xi, const double *coef_eta, int *isrc);

/* Source amplitude table: ampl[it-1] = STF_get(it*dt - tdelay)*ampli for it = 1..nt
   (SRC/src_gen.f90:290-317, SRC/stf_gen.f90:106-127); steps beyond nt use 0. */
int sem2d_set_stf(sem2d_solver *h, int isrc, int nt, const double *ampl);

/* Receivers at GLL nodes: replaces REC_store with AtNode=T (SRC/receivers.f90:309-344).
   field: 0 = displacement, 1 = velocity, 2 = acceleration.  A sample is taken at the
   current step when registered (step 0) and after every step it with it % isamp == 0;
   nsamp samples are kept. */
int sem2d_set_receivers(sem2d_solver *h, int nrec, const int *iglob, int isamp, int field,
                        int nsamp);

/* Dynamic fault: replaces BC_DYNFLT_apply (SRC/bc_dynflt.f90:569-689) and the
   record part of BC_DYNFLT_write (SRC/bc_dynflt.f90:751-778).  nsamp records are
   kept; a record is taken at registration when the current step >= oit. */
int sem2d_set_fault(sem2d_solver *h, const sem2d_fault_desc *f, int nsamp);

/* Dirichlet components of a boundary: replaces BC_DIRNEU_apply with h or v = 'D'
   (SRC/bc_dirneu.f90:148-169): the force (MxA) of the n nodes (1-based) is set to zero in
   the components of comp (1: component 1, 2: component 2, 3: both; SH: 1).  Boundary
   conditions other than absorbing ones act in the order they are registered
   (SRC/bc_gen.f90:273-281): a Dirichlet boundary registered after sem2d_set_fault acts
   after the fault solve, one registered before it acts before. */
int sem2d_set_dirichlet(sem2d_solver *h, int n, const int *node, int comp);

/* Kelvin-Voigt damping: eta [nelem][ngll][ngll] per element point, already scaled by dt
   when ETAxDT (matwrk%eta of MAT_KV_init_elem_work, SRC/mat_kelvin_voigt.f90:121-135),
   0 outside Kelvin-Voigt elements; the element force becomes -K (d + eta v)
   (SRC/mat_gen.f90:435).  Call right after sem2d_create. */
int sem2d_set_kelvin_voigt(sem2d_solver *h, const double *eta);

/* 2.5D crustal-plane term (Lapusta-Rice; &GENERAL W): beta [nelem][ngll][ngll] per element
   point (matwrk%beta of MAT_ELAST_init_25D, SRC/mat_elastic.f90:363-384); the element force
   becomes -K d - beta d, d including the Kelvin-Voigt term (MAT_ELAST_add_25D_f,
   SRC/mat_elastic.f90:447-459, called from MAT_Fint, SRC/mat_gen.f90:440).  Call right
   after sem2d_create. */
int sem2d_set_25d(sem2d_solver *h, const double *beta);

/* Receivers inside elements (REC_LINE AtNode=F; REC_posit/REC_store,
   SRC/receivers.f90:274-285,326-341): element elem[r] (1-based) and its 25 weights
   interp[r][k] = h_i(xi) h_j(eta), k = (j-1)*5 + i, as SE_init_interpol
   (SRC/spec_grid.f90:385-406) computes them; a sample is sum_k interp[r][k] *
   field(ibool(k, elem[r])).  Replaces any receivers set before. */
int sem2d_set_receivers_interp(sem2d_solver *h, int nrec, const int *elem, const double *interp,
                               int isamp, int field, int nsamp);

/* Advance nsteps Newmark steps: replaces the loop body of SRC/main.f90:51-99
   (solve -> solve_Newmark, SRC/solver.f90:42-84, plus REC_store and BC_write records).
   Returns once the device has finished the nsteps steps (their mean time is then
   available from sem2d_last_step_ms); a nonzero code reports a launch failure or the
   Newton solver's iteration overflow (the reference's IO_abort in nr_solver). */
int sem2d_step(sem2d_solver *h, int nsteps);

/* Fields [ndof][npoin] (any pointer may be NULL). */
int sem2d_fetch_fields(sem2d_solver *h, double *d, double *v, double *a);
int sem2d_set_fields(sem2d_solver *h, const double *d, const double *v, const double *a);

/* Seismograms [nsamp][ndof][nrec] in FP64 (the reference stores float32). */
int sem2d_fetch_seismo(sem2d_solver *h, double *out, int *nsamp_taken);

/* Fault records [nsamp][6][nout]: slip, slip rate, shear traction, normal traction,
   friction coefficient, stick traction (the six lines of SRC/bc_dynflt.f90:762-767)
   with nout = (oixn-oix1)/oixd + 1. */
int sem2d_fetch_fault(sem2d_solver *h, double *out, int *nsamp_taken);

/* Per-side fault records (BC_DYNFLT osides = T; export_side, SRC/bc_dynflt.f90:769-812):
   after sem2d_set_fault and before the first step, on = 1 makes every fault record also
   keep [4*ndof][nout] values: the displacement of side 1, of side 2, the velocity of side
   1, of side 2, each in the fault frame (t, n) for P-SV (rotate, SRC/bc_dynflt.f90:722-739)
   and side 2 = -(side 1) on a fault with one side.  sem2d_fetch_fault_sides returns them
   as [nsamp][4*ndof][nout], the rows the reference appends to each FltXX_sem2d.dat record. */
int sem2d_set_fault_sides(sem2d_solver *h, int on);

/* Time-weakening friction (BC_DYNFLT_TWF, SRC/bc_dynflt_twf.f90:129-183; the forced
   nucleation of BC_DYNFLT friction = 'SWF','TWF'): after sem2d_set_fault and before the
   first step.  kind 1..3, par = MuS, MuD, Mu0, X, Z, V, L, T, Dc, coord [2][npoin] the fault
   node coordinates (bc%coord).  The friction coefficient becomes the minimum of the
   slip-weakening or rate-and-state one and twf_mu(time = it*dt) (SRC/bc_dynflt.f90:404-413,
   632,653), or twf_mu alone with law 3.  With kind 3 the reference leaves the coefficient
   of nodes beyond V*T unset; here it is huge(1d0), i.e. no weakening there. */
int sem2d_set_fault_twf(sem2d_solver *h, int kind, const double *par, const double *coord);
int sem2d_fetch_fault_sides(sem2d_solver *h, double *out, int *nsamp_taken);

/* Normal stress response of the fault (BC_DYNFLT_NOR; normal_init / normal_update,
   SRC/bc_dynflt_normal.f90:104-140): after sem2d_set_fault and before the first step.
   kind 0: the friction sees the initial normal stress throughout; 1: Coulomb (the trial
   normal traction, the default); 2: Prakash-Clifton with a time scale, sigma = Tn +
   coef (sigma - Tn), coef = exp(-dt/T); 3: with a length scale, sigma = Tn +
   exp(-(|V| + v) coef) (sigma - Tn), coef = dt/L, V the trial slip rate. */
int sem2d_set_fault_normal(sem2d_solver *h, int kind, double coef, double v);

/* Incident plane wave (SRC_DEF mechanism = 'WAVE', &SRC_WAVE) through the absorbing
   boundary registered last by sem2d_set_abso (BC_ABSO_apply with an incident wave,
   SRC/bc_abso.f90:300-306; SO_WAVE_get_VT, SRC/src_gen.f90:322-339).  xz [2][n] the
   boundary nodes' coordinates, vdir and tdir [ndof][n] the incident velocity and traction
   for unit amplitude (SRC_WAVE_get_VT, SRC/src_wave.f90:134-171, with the node normals),
   B [n] the boundary weights, par = s1, s2, x0, z0 (slowness and phase origin), tdelay,
   and the Ricker f0, onset and ampli.  Each step the force at node k becomes
   f - C (v - a Vdir) + B a Tdir, a = ricker(t - (x - x0) s1 - (z - z0) s2 - tdelay), t the
   boundary conditions' time.  The problem steps on the general element path. */
int sem2d_set_abso_wave(sem2d_solver *h, int n, const double *xz, const double *vdir, const double *tdir,
                        const double *B, const double *par);

/* The absorbing boundary registered last ends on a periodic boundary at both ends
   (bc%periodic, SRC/bc_abso.f90:142): after sem2d_set_abso_stacey, its two end nodes
   each take the Stacey terms of both (SRC/bc_abso.f90:325-328).  General element path. */
int sem2d_set_abso_periodic(sem2d_solver *h);

/* Linear slip fault (BC_DEF kind = 'LISFLT'; BC_LSF_init / BC_LSF_apply,
   SRC/bc_lsf.f90:101-186): n node pairs (1-based) of the two sides, the side-1 normals
   [2][n], B, W1 = 1/(B (1 + M1/M2)), W2 = 1/(B (M2/M1 + 1)) [ndof][n], the tangential and
   normal stiffness kt, kn (DBL_MAX: welded).  After the absorbing boundaries, each step
   T = W2 f2 - W1 f1 and dD = d2 - d1 in the fault frame, T_t = kt dD_t, T_n = kn dD_n,
   f1 += B T, f2 -= B T.  Not combined with sem2d_set_fault or Dirichlet boundaries; the
   problem steps on the general element path. */
int sem2d_set_lsf(sem2d_solver *h, int n, const int *node1, const int *node2, const double *nrm,
                  const double *B, const double *W1, const double *W2, double kt, double kn);

/* HHT-alpha time scheme (TIME kind = 'HHT-alpha'; solve_HHT_alpha, SRC/solver.f90:89-131):
   the descriptor carries the scheme's beta and gamma (gamma = 3/2 - alpha, beta from rho,
   SRC/time.f90:219-235).  Each step predicts d and v as Newmark does, forms the fields at
   n + alpha, da = alpha d + (1 - alpha) d_old and va likewise, and computes the element
   forces (Kelvin-Voigt with va) and the absorbing terms from them; sources and boundary
   conditions see t_alpha = it dt + (alpha - 1) dt (the caller's amplitude tables are taken
   at t_alpha).  Call it right after sem2d_create, before sem2d_set_kelvin_voigt and
   sem2d_set_25d: the problem then steps on the general element path. */
int sem2d_set_hht(sem2d_solver *h, double alpha);

/* Periodic boundaries (BC_DEF kind = 'PERIOD'; BC_PERIO_init / BC_PERIO_apply,
   SRC/bc_periodic.f90:45-104): n node pairs (1-based), the master boundary's nodes and the
   slave boundary's matching ones.  Each step, after the sources and before every other
   boundary condition (SRC/bc_gen.f90:273-275), the force of each master node becomes the
   sum of both and the slave node takes it.  The mass given to sem2d_create is the
   assembled one (master and slave equal).  Call it right after sem2d_create, before
   sem2d_set_kelvin_voigt and sem2d_set_25d: the problem then steps on the general element
   path. */
int sem2d_set_periodic(sem2d_solver *h, int n, const int *master, const int *slave);

/* Fault potency rows [nrows][2*(ndof+1)], one per step from it = 0: P-SV p11, p22, p12
   of the displacement jump then of the velocity jump, SH p13, p23 of each -- the
   FltXX_potency_sem2d.tab lines of BC_DYNFLT_write (SRC/bc_dynflt.f90:757,832-855). */
int sem2d_fetch_potency(sem2d_solver *h, double *out, int *nrows);

/* Current fault state of all pairs: D, V [ndof][npoin], T, Tstick [2][npoin], MU [npoin]. */
int sem2d_fetch_fault_state(sem2d_solver *h, double *D, double *V, double *T, double *MU,
                            double *Tstick);

/* f = -K d assembled over all elements: compute_Fint (SRC/solver.f90:273-320) with
   MAT_ELAST_f (SRC/mat_elastic.f90:396-426).  d and f are [ndof][npoin] host arrays. */
int sem2d_compute_fint(sem2d_solver *h, const double *d, double *f);

/* sem2d_compute_fint without device allocation: no counterpart in the reference.  For
   callers that cannot allocate device memory (code under HIP graph capture, serving
   loops, harnesses that forbid hipMalloc).  Same arguments, kernels and results as
   sem2d_compute_fint; its staging buffers are a file-scope device array of 64 MiB per
   GPU (held until the process ends), so ndof*npoin is limited to 4,194,304 values
   (error 3 beyond).  Calls on one GPU share that array: calls that overlap in time
   (two host threads, two handles) race. */
int sem2d_compute_fint_noalloc(sem2d_solver *h, const double *d, double *f);

/* Time scheme (the dispatch of solve, SRC/solver.f90:20-37): kind 0 is the Newmark
   scheme of the descriptor's beta and gamma (solve_Newmark, SRC/solver.f90:42-84, the
   default); kind 1 the central-difference scheme (solve_leapfrog, SRC/solver.f90:140-160:
   d = d + dt v, forces, a = f/M, v = v + dt a, with v the mid-step velocity), which needs
   beta = 0, gamma = 1/2 in the descriptor and a fault descriptor with CoefA2V = dt
   (TIME_getCoefA2V, SRC/time.f90:445-458).  Call before the first step. */
int sem2d_set_time_scheme(sem2d_solver *h, int kind);

/* Progress information of the time loop (SRC/main.f90:45,73-76: every ItInfo steps the
   reference prints it, it*dt, maxval(abs(veloc)), maxval(abs(displ))).  sem2d_set_itinfo
   makes sem2d_step record such a row on the device every itinfo steps (0: off, the
   default); each row also counts the values of d, v, a that are not finite, and
   sem2d_step returns 21 ("non-finite field values at step N") once a row has found one
   (the reference has no such check; its progress line would show NaN).  The error stays
   until the handle is destroyed. */
int sem2d_set_itinfo(sem2d_solver *h, int itinfo);
/* rows [nrows][4]: it, time, vmax, dmax */
int sem2d_fetch_info(sem2d_solver *h, double *rows, int *nrows);
/* the same quantities for the current fields, now; nonfinite = count over d, v, a */
int sem2d_field_info(sem2d_solver *h, double *vmax, double *dmax, int *nonfinite);

/* Launch configuration of the structured-grid tile kernel (no counterpart in the
   reference, whose loop is serial): threads per workgroup (256 = 10-element batches,
   512 = 20-element batches) and persistent workgroups per CU (0: one workgroup per
   tile; < 0: automatic, the persistent kernel with 2 (FP64) or 3 (FP32) workgroups
   per CU once there are >= 4 tiles per workgroup).  Results do not depend on it (same
   summation order). */
int sem2d_set_tile_config(sem2d_solver *h, int threads, int wg_per_cu);

/* Current step index (number of steps taken). */
int sem2d_get_step(sem2d_solver *h, int *it);

/* Average duration in ms of the kernels launched by the last sem2d_step call, measured
   with HIP events on the handle's stream (force+step kernels only). */
int sem2d_last_step_ms(sem2d_solver *h, double *ms_per_step);

const char *sem2d_last_error(sem2d_solver *h);
void sem2d_destroy(sem2d_solver *h);

/* ------------------------------------------------------------------------------
   Adjoint of the time loop (no counterpart in the reference, which has no
   differentiation).  Structured P-SV grids with compact (lambda, mu) coefficients,
   beta = 0, no fault.  J is any function of the recorded seismograms.

   sem2d_adjoint_init      plan the storage of the forward states for nsteps steps.
                           seg > 0: sqrt(N)-style checkpointing, the state before every
                           segment of seg steps plus seg+1 replay slots.  seg <= 0: the
                           displacement history of every step (FP32 fields: with the
                           acceleration) if it fits in free device memory, else
                           checkpoints with seg = ceil(sqrt(N))
   sem2d_adjoint_forward   from the zero state, nsteps steps storing the checkpoints or
                           the history; seismograms are recorded as by sem2d_step
   sem2d_adjoint_backward  gseis = dJ/d(seismograms) [nsamp][ndof][nrec]; runs the
                           adjoint (transposed) steps backwards, replaying each segment
                           from its checkpoint first (checkpoint plan) or reading the
                           history (history plan: a_n and the predicted velocity are
                           differences of the stored displacements)
   sem2d_adjoint_fetch     the sensitivities: glam, gmu [nelem] = dJ/d(lambda), dJ/d(mu)
                           per element (summed over its GLL points), gmass [npoin] =
                           dJ/dM (assembled mass, both components), gabso [ndof][npoin] =
                           dJ/dC at absorbing-boundary nodes (0 elsewhere), C the
                           Clayton-Engquist coefficient including its share of the
                           modified mass
   ------------------------------------------------------------------------------ */
int sem2d_adjoint_init(sem2d_solver *h, int nsteps, int seg);
int sem2d_adjoint_forward(sem2d_solver *h);
int sem2d_adjoint_backward(sem2d_solver *h, const double *gseis);
int sem2d_adjoint_fetch(sem2d_solver *h, double *glam, double *gmu, double *gmass, double *gabso);
/* device time (ms) of the last forward and backward sweeps */
int sem2d_adjoint_timing(sem2d_solver *h, double *fwd_ms, double *bwd_ms);
/* the plan sem2d_adjoint_init chose: history = 1 for the history plan, seg = the
   checkpoint segment length (nsteps with the history plan) */
int sem2d_adjoint_plan(sem2d_solver *h, int *history, int *seg);
/* The history plan's device block outlives its solver: it is kept (one block per
   process) for the next history plan that fits in it, since releasing and mapping a
   block of that size costs seconds.  This returns the kept block to the device. */
int sem2d_adjoint_release(void);

/* ------------------------------------------------------------------------------
   Batched ensemble: many independent scenarios advanced together.

   The reference runs one scenario per process (SRC/main.f90:51-99, one time loop per
   problem_type); a batch replaces N such time loops.  Each member is an ordinary
   handle, fully set up (sem2d_create, boundaries, sources, receivers, fault) and not
   yet stepped.  The members must be MESH_CART grids (sem2d_desc.nx > 0) with the same
   ndof, element kernel, coefficient mode, bulk precision, dt and Newmark parameters;
   with faults, the same friction law and kind, opening and healing flags and output
   times.  Mesh sizes, materials, boundaries, sources, receivers and fault parameters
   may differ.  The batch copies the members' state at creation and then evolves on
   its own; every member computes exactly what its own sem2d_step would (bitwise).
   The members must outlive the batch (their node maps are used by the fetches).
   ------------------------------------------------------------------------------ */
typedef struct sem2d_batch sem2d_batch;

int sem2d_batch_create(int n, sem2d_solver *const *members, sem2d_batch **out);

/* Advance every member by nsteps.  The loop runs as replays of a captured hipGraph of
   a chunk of up to 64 steps (the step index is kept in device memory). */
int sem2d_batch_step(sem2d_batch *b, int nsteps);

/* Autotuner: time trial_steps steps of each tile-kernel launch configuration
   (threads per workgroup = elements per batch x 25, workgroups per CU, persistent or
   one workgroup per tile), keep the fastest and restore the state.  *choice receives
   the index of the chosen configuration. */
int sem2d_batch_tune(sem2d_batch *b, int trial_steps, int *choice);

/* Set the tile-kernel launch configuration explicitly (as sem2d_set_tile_config;
   0 <= wg_per_cu <= 4). */
int sem2d_batch_set_config(sem2d_batch *b, int threads, int wg_per_cu);

/* Current tile-kernel launch configuration and total tile count. */
int sem2d_batch_config(sem2d_batch *b, int *threads, int *workgroups, int *ntiles);

/* Per-member outputs, in the layouts of the single-scenario fetches. */
int sem2d_batch_fetch_fields(sem2d_batch *b, int member, double *d, double *v, double *a);
int sem2d_batch_fetch_seismo(sem2d_batch *b, int member, double *out, int *nsamp_taken);
int sem2d_batch_fetch_fault(sem2d_batch *b, int member, double *out, int *nsamp_taken);
int sem2d_batch_fetch_fault_state(sem2d_batch *b, int member, double *D, double *V, double *T,
                                  double *MU, double *Tstick);
/* Potency rows of a member, as sem2d_fetch_potency (the FltXX_potency_sem2d.tab lines of
   BC_DYNFLT_write, SRC/bc_dynflt.f90:757,832-855): the member's row of step 0, then one row
   per batched step. */
int sem2d_batch_fetch_potency(sem2d_batch *b, int member, double *out, int *nrows);
int sem2d_batch_get_step(sem2d_batch *b, int *it);
/* ms per step of the last sem2d_batch_step (HIP events on the batch stream) */
int sem2d_batch_last_step_ms(sem2d_batch *b, double *ms_per_step);
const char *sem2d_batch_last_error(sem2d_batch *b);
void sem2d_batch_destroy(sem2d_batch *b);

#ifdef __cplusplus
}
#endif
#endif /* SEM2D_H */
