/*
 * sem2d_oracle.c -- CPU restatement of the SEM2DPACK explicit time step.
 *
 * TEST INFRASTRUCTURE ONLY: this is the oracle against which the HIP library is
 * checked.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py load it; the product never does.
 *
 * It implements the C ABI of include/sem2d.h under the prefix oracle_, in plain
 * sequential C that follows the reference's Fortran statement by statement and in
 * the same floating-point operation order (compile with -ffp-contract=off):
 *
 *   oracle_step          solve_Newmark, solve_HHT_alpha, solve_leapfrog
 *                        SRC/solver.f90:42-84,89-131,140-160
 *   fint()               compute_Fint         SRC/solver.f90:273-320 (element order)
 *     gather/scatter     FIELD_get_elem_sub / FIELD_add_elem_2  SRC/fields.f90:113-189
 *     kd()               ELAST_KD2_PSV/SH (NGLL 5), ELAST_KD1_PSV/SH (other NGLL),
 *                        My_MATMUL, mxm   SRC/mat_elastic.f90:464-518,589-799, SRC/mxmlib.f90
 *   sources              SO_add, FORCE_add, SRC_MOMENT_add  SRC/src_gen.f90:290-317,
 *                        SRC/src_force.f90:77-90, SRC/src_moment.f90:217-235
 *   abso_apply           BC_ABSO_apply (flat and not flat; Stacey P-SV)  SRC/bc_abso.f90:286-336
 *   fault_apply          BC_DYNFLT_apply      SRC/bc_dynflt.f90:569-739
 *     swf                swf_mu, swf_update_state  SRC/bc_dynflt_swf.f90:142-194
 *     rsf                rsf_solver, rsf_update_theta, rsf_update_V, nr_solver,
 *                        nr_fric_func_tau, rsf_mu  SRC/bc_dynflt_rsf.f90:165-568
 *     normal             normal_update kinds 0-3  SRC/bc_dynflt_normal.f90:122-140
 *   periodic             BC_PERIO_apply       SRC/bc_periodic.f90:78-90, SRC/bc_gen.f90:273-275
 *   receivers            REC_store (AtNode)   SRC/receivers.f90:309-344
 *   fault records        BC_DYNFLT_write      SRC/bc_dynflt.f90:751-778
 *   potency              BC_DYNFLT_potency    SRC/bc_dynflt.f90:832-855
 *   Kelvin-Voigt         MAT_KV_add_etav      SRC/mat_kelvin_voigt.f90:139-150
 *
 * Known deviation: in the rate-and-state branch the reference leaves the local
 * Tstick array unset (SRC/bc_dynflt.f90:620-636 never assigns it); here Tstick is
 * the stick traction minus T0 for both friction laws.
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/sem2d.h"

/* Perturbed builds (test infrastructure for measuring the conditioning of the
   reference arithmetic; the plain build defines none of these):
   ORACLE_ULP_PERTURB=+1 / -1  every exp/log/sinh/cosh/asinh result moved by one ulp,
                               in opposite directions for the two signs;
   ORACLE_ULP_RANDOM=seed      every exp/log/sinh/cosh/asinh/pow result moved by -1, 0 or
                               +1 ulp, the direction drawn per call from a seeded
                               xorshift sequence (the pattern of a libm that agrees
                               with the C library to within one ulp either way);
   ORACLE_ORDER=1              element forces assembled in reverse element order;
   ORACLE_ORDER=2              element forces assembled colour by colour (greedy
                               first-fit colouring of the elements in element order,
                               each colour in element order), the summation order of a
                               colour-scheduled GPU assembly;
   a build with -ffp-contract=fast -mfma contracts products and sums into FMAs. */
#ifdef ORACLE_ULP_PERTURB
#define ULP_S ((double)(ORACLE_ULP_PERTURB) * 0x1p-52)
static double ulp_exp(double x) { return exp(x) * (1.0 + ULP_S); }
static double ulp_log(double x) { return log(x) * (1.0 - ULP_S); }
static double ulp_sinh(double x) { return sinh(x) * (1.0 + ULP_S); }
static double ulp_cosh(double x) { return cosh(x) * (1.0 - ULP_S); }
static double ulp_asinh(double x) { return asinh(x) * (1.0 + ULP_S); }
#define exp ulp_exp
#define log ulp_log
#define sinh ulp_sinh
#define cosh ulp_cosh
#define asinh ulp_asinh
#endif
#ifdef ORACLE_ULP_RANDOM
static unsigned long long ulp_state = 0x9E3779B97F4A7C15ull * (unsigned long long)(ORACLE_ULP_RANDOM);
static double ulp_rand(double y) {
  ulp_state ^= ulp_state << 13;
  ulp_state ^= ulp_state >> 7;
  ulp_state ^= ulp_state << 17;
  const int s = (int)(ulp_state % 3ull) - 1;
  return y * (1.0 + s * 0x1p-52);
}
static double ulp_exp(double x) { return ulp_rand(exp(x)); }
static double ulp_log(double x) { return ulp_rand(log(x)); }
static double ulp_sinh(double x) { return ulp_rand(sinh(x)); }
static double ulp_cosh(double x) { return ulp_rand(cosh(x)); }
static double ulp_asinh(double x) { return ulp_rand(asinh(x)); }
static double ulp_pow(double x, double y) { return ulp_rand(pow(x, y)); }
#define exp ulp_exp
#define log ulp_log
#define sinh ulp_sinh
#define cosh ulp_cosh
#define asinh ulp_asinh
#define pow ulp_pow
#endif
#ifndef ORACLE_ORDER
#define ORACLE_ORDER 0
#endif

#define NGMAX 10 /* largest NGLL (the reference's mxm library covers 2..10) */

typedef struct {
  int n;
  int *node;    /* 0-based */
  double *C;    /* [ndof][n] */
  double *nrm;  /* [2][n] node normals of a boundary that is not flat, NULL if flat */
  int nbe;      /* Stacey: boundary elements (0: none) */
  int perends;  /* the boundary ends on periodic boundaries: Stacey end terms summed */
  double *wxz, *wV, *wT, *wB; /* incident plane wave: [2][n], [ndof][n], [ndof][n], [n] */
  double wpar[8];             /* s1, s2, x0, z0, tdelay, f0, t0, ampli */
  int *enode;   /* [nbe][ngll] boundary-local node of each edge point */
  double *K;    /* [nbe][ngll][2] */
} abso_t;

typedef struct {
  int kind; /* 1 force, 2 moment */
  int iglob;
  double dir[2];
  int nel;
  int *ixi, *ieta;       /* [nel][ngll] 0-based */
  double *cxi, *ceta;    /* [nel][ndof][ngll] */
  int nt;
  double *ampl;
} source_t;

typedef struct {
  int np, ndof;
  int *n1, *n2; /* 0-based */
  double *nrm;  /* [2][np] */
  double *B, *invM1, *invM2, *Z; /* [ndof][np] */
  double *T0, *coh;
  double a2v, a2d;
  int opening, law, kind, healing;
  double *p[6];
  double *theta, *Tc, *coeft;
  double *D, *V, *T, *Tstick, *MU, *sigma;
  int oit, oitd, oix1, oixn, oixd, nout;
  int nsamp, taken;
  double *rec; /* [nsamp][6][nout] */
  int twf;       /* time-weakening friction kind (0: none), law 3: time weakening alone */
  double tw[9];  /* MuS, MuD, Mu0, X, Z, V, L, T, Dc */
  double *xz;    /* [2][np] fault node coordinates (bc%coord) */
  int nor_kind;      /* normal stress response: 0 cohesion, 1 Coulomb, 2-3 Prakash-Clifton */
  double nor_coef;   /* kind 2: exp(-dt/T); kind 3: dt/L */
  double nor_v;      /* kind 3: V */
  int osides;   /* per-side records (osides = T) */
  double *srec; /* [nsamp][4*ndof][nout]: d of side 1, side 2, v of side 1, side 2 */
  int pot_rows, pot_cap;
  double *pot; /* [rows][2*(ndof+1)] potency of d then v, one row per step from it = 0 */
} fault_t;

struct oracle_solver {
  int ndof, nelem, npoin, nelast, shared;
  int *ibool; /* 0-based [nelem][ngll][ngll] */
  double *coefs;
  int ng, npt;      /* NGLL, NGLL^2 */
  double H[NGMAX][NGMAX]; /* H[i][j] = h'_i(x_j) */
  double *rmass;
  double dt, beta, gamma;
  int leapfrog;     /* 1: central-difference scheme (SRC/solver.f90:140-160) */
  int hht;          /* 1: HHT-alpha (SRC/solver.f90:89-131) */
  double alpha;
  double *da, *va;  /* displ_alpha, veloc_alpha */
  double *d, *v, *a;
  int nabso;
  abso_t abso[16];
  int ndir;            /* Dirichlet boundaries (BC_DIRNEU 'D' components) */
  int *dir_node[16];   /* 0-based nodes */
  int dir_n[16], dir_comp[16], dir_after_fault[16];
  int nsrc;
  source_t src[64];
  int nrec, isamp, field, nsamp, taken;
  int *rnode;        /* [nrec], or [nrec][25] element nodes when interpolated */
  double *rw;        /* [nrec][25] interpolation weights, NULL for receivers at nodes */
  int nlsf;          /* linear slip fault (BC_LSF): node pairs (0-based) and arrays */
  int *lsf1, *lsf2;
  double *lsf_nrm, *lsf_B, *lsf_W1, *lsf_W2, lsf_kt, lsf_kn;
  int nper;          /* periodic node pairs (BC_PERIO): master, slave (0-based) */
  int *per_m, *per_s;
  double *kv_eta;    /* [nelem][25] Kelvin-Voigt eta (dt-scaled), NULL if none */
  double *beta25;    /* [nelem][25] 2.5D coefficients (matwrk%beta), NULL if none */
  int *eorder;       /* assembly order of the elements */
  double *floc;      /* [nelem][ndof][ngll][ngll] element forces of the current step */
  double *seis;
  fault_t *flt;
  int it;
  int itinfo, info_rows, info_cap, first_bad; /* progress rows every itinfo steps */
  double *info;                               /* [rows][4]: it, time, vmax, dmax */
  char err[256];
};
typedef struct oracle_solver oracle_solver;

static int fail(oracle_solver *h, int code, const char *msg) {
  if (h) snprintf(h->err, sizeof h->err, "%s", msg);
  return code;
}

static double *dup_d(const double *src, size_t n) {
  double *p = (double *)malloc(n * sizeof(double) + 8);
  if (src) memcpy(p, src, n * sizeof(double)); else memset(p, 0, n * sizeof(double));
  return p;
}

int oracle_create(const sem2d_desc *s, oracle_solver **out) {
  if (!s || !out) return 1;
  if (s->ngll < 3 || s->ngll > NGMAX) return 2;
  oracle_solver *h = (oracle_solver *)calloc(1, sizeof *h);
  h->ng = s->ngll;
  h->npt = s->ngll * s->ngll;
  h->ndof = s->ndof; h->nelem = s->nelem; h->npoin = s->npoin; h->nelast = s->nelast;
  if (s->coef_mode == 2) { free(h); return 2; } /* the oracle uses the reference's a() arrays */
  h->shared = s->coef_mode == 1;
  h->ibool = (int *)malloc(sizeof(int) * (size_t)s->nelem * h->npt);
  for (size_t k = 0; k < (size_t)s->nelem * h->npt; ++k) h->ibool[k] = s->ibool[k] - 1;
  size_t nc = (size_t)(h->shared ? 1 : s->nelem) * s->nelast * h->npt;
  h->coefs = dup_d(s->coefs, nc);
  for (int j = 0; j < h->ng; ++j)
    for (int i = 0; i < h->ng; ++i) h->H[i][j] = s->hprime[j * h->ng + i];
  size_t nf = (size_t)s->ndof * s->npoin;
  h->rmass = dup_d(s->rmass, nf);
  h->dt = s->dt; h->beta = s->beta; h->gamma = s->gamma;
  h->d = dup_d(NULL, nf); h->v = dup_d(NULL, nf); h->a = dup_d(NULL, nf);
  *out = h;
  return 0;
}

int oracle_set_abso(oracle_solver *h, int n, const int *node, const double *C) {
  if (h->nabso >= 16) return fail(h, 3, "too many absorbing boundaries");
  abso_t *b = &h->abso[h->nabso++];
  b->n = n;
  b->node = (int *)malloc(sizeof(int) * n);
  for (int k = 0; k < n; ++k) b->node[k] = node[k] - 1;
  b->C = dup_d(C, (size_t)h->ndof * n);
  return 0;
}

/* BC_DIRNEU with Dirichlet components (SRC/bc_dirneu.f90:148-169): MxA(node, c) = 0 for the
   components in comp (bit 0: component 1, bit 1: component 2), in registration order
   among the boundary conditions other than absorbing ones (the fault's place included) */
int oracle_set_dirichlet(oracle_solver *h, int n, const int *node, int comp) {
  if (h->ndir >= 16) return fail(h, 3, "too many Dirichlet boundaries");
  if (n <= 0 || !node || comp < 1 || comp > 3) return fail(h, 3, "set_dirichlet: bad arguments");
  int b = h->ndir++;
  h->dir_node[b] = (int *)malloc(sizeof(int) * n);
  for (int k = 0; k < n; ++k) {
    if (node[k] < 1 || node[k] > h->npoin) return fail(h, 3, "set_dirichlet: node out of range");
    h->dir_node[b][k] = node[k] - 1;
  }
  h->dir_n[b] = n;
  h->dir_comp[b] = comp;
  h->dir_after_fault[b] = h->flt != NULL;
  return 0;
}

static void dirichlet_apply(oracle_solver *h, int b, double *MxA) {
  for (int c = 0; c < h->ndof; ++c)
    if (h->dir_comp[b] >> c & 1)
      for (int k = 0; k < h->dir_n[b]; ++k) MxA[(size_t)c * h->npoin + h->dir_node[b][k]] = 0.0;
}

int oracle_set_abso_normals(oracle_solver *h, int n, const double *nrm) {
  if (h->nabso < 1) return fail(h, 3, "set_abso_normals: no absorbing boundary");
  if (h->ndof != 2) return fail(h, 3, "set_abso_normals: P-SV only");
  abso_t *b = &h->abso[h->nabso - 1];
  if (n != b->n || !nrm) return fail(h, 3, "set_abso_normals: size differs from the boundary's");
  b->nrm = dup_d(nrm, (size_t)2 * n);
  return 0;
}

int oracle_set_abso_stacey(oracle_solver *h, int nbe, const int *enode, const double *K) {
  if (h->nabso < 1) return fail(h, 3, "set_abso_stacey: no absorbing boundary");
  if (h->ndof != 2) return fail(h, 3, "set_abso_stacey: P-SV only");
  abso_t *b = &h->abso[h->nabso - 1];
  if (nbe < 1 || !enode || !K) return fail(h, 3, "set_abso_stacey: empty");
  for (int k = 0; k < nbe * h->ng; ++k)
    if (enode[k] < 0 || enode[k] >= b->n) return fail(h, 3, "set_abso_stacey: node out of range");
  b->nbe = nbe;
  b->enode = (int *)malloc(sizeof(int) * (size_t)nbe * h->ng);
  memcpy(b->enode, enode, sizeof(int) * (size_t)nbe * h->ng);
  b->K = dup_d(K, (size_t)nbe * h->ng * 2);
  return 0;
}

int oracle_set_force(oracle_solver *h, int iglob, const double *dir, int *isrc) {
  source_t *s = &h->src[h->nsrc];
  memset(s, 0, sizeof *s);
  s->kind = 1; s->iglob = iglob - 1;
  if (dir) { s->dir[0] = dir[0]; s->dir[1] = dir[1]; }
  if (isrc) *isrc = h->nsrc;
  h->nsrc++;
  return 0;
}

int oracle_set_moment(oracle_solver *h, int nel, const int *ixi, const int *ieta,
                      const double *cxi, const double *ceta, int *isrc) {
  source_t *s = &h->src[h->nsrc];
  memset(s, 0, sizeof *s);
  s->kind = 2; s->nel = nel;
  s->ixi = (int *)malloc(sizeof(int) * nel * h->ng);
  s->ieta = (int *)malloc(sizeof(int) * nel * h->ng);
  for (int k = 0; k < nel * h->ng; ++k) { s->ixi[k] = ixi[k] - 1; s->ieta[k] = ieta[k] - 1; }
  s->cxi = dup_d(cxi, (size_t)nel * h->ndof * h->ng);
  s->ceta = dup_d(ceta, (size_t)nel * h->ndof * h->ng);
  if (isrc) *isrc = h->nsrc;
  h->nsrc++;
  return 0;
}

int oracle_set_stf(oracle_solver *h, int isrc, int nt, const double *ampl) {
  if (isrc < 0 || isrc >= h->nsrc) return fail(h, 3, "bad source index");
  h->src[isrc].nt = nt;
  h->src[isrc].ampl = dup_d(ampl, nt);
  return 0;
}

static const double *rec_field(oracle_solver *h) {
  return h->field == 0 ? h->d : (h->field == 1 ? h->v : h->a);
}

static void rec_store(oracle_solver *h) {
  if (!h->nrec || h->it % h->isamp) return;
  int s = h->it / h->isamp;
  if (s >= h->nsamp) return;
  const double *f = rec_field(h);
  for (int c = 0; c < h->ndof; ++c)
    for (int r = 0; r < h->nrec; ++r) {
      double v;
      if (h->rw) { /* matmul(interp, vloc): SRC/receivers.f90:329-339 */
        v = 0.0;
        for (int k = 0; k < h->npt; ++k)
          v += h->rw[(size_t)r * h->npt + k] * f[(size_t)c * h->npoin + h->rnode[(size_t)r * h->npt + k]];
      } else {
        v = f[(size_t)c * h->npoin + h->rnode[r]];
      }
      h->seis[((size_t)s * h->ndof + c) * h->nrec + r] = v;
    }
  if (s + 1 > h->taken) h->taken = s + 1;
}

int oracle_set_receivers_interp(oracle_solver *h, int nrec, const int *elem, const double *interp, int isamp,
                                int field, int nsamp) {
  h->nrec = nrec; h->isamp = isamp; h->field = field; h->nsamp = nsamp;
  h->rnode = (int *)malloc(sizeof(int) * nrec * h->npt);
  for (int r = 0; r < nrec; ++r)
    for (int k = 0; k < h->npt; ++k) h->rnode[r * h->npt + k] = h->ibool[(size_t)(elem[r] - 1) * h->npt + k];
  h->rw = dup_d(interp, (size_t)nrec * h->npt);
  h->seis = dup_d(NULL, (size_t)nsamp * h->ndof * nrec);
  h->taken = 0;
  rec_store(h);
  return 0;
}

int oracle_set_receivers(oracle_solver *h, int nrec, const int *iglob, int isamp, int field,
                         int nsamp) {
  h->nrec = nrec; h->isamp = isamp; h->field = field; h->nsamp = nsamp;
  h->rnode = (int *)malloc(sizeof(int) * nrec);
  for (int r = 0; r < nrec; ++r) h->rnode[r] = iglob[r] - 1;
  h->seis = dup_d(NULL, (size_t)nsamp * h->ndof * nrec);
  h->taken = 0;
  rec_store(h);
  return 0;
}

/* ---------------- friction ---------------- */
/* twf_mu (SRC/bc_dynflt_twf.f90:129-183) at node k, time t, slip d; kind 3 leaves
   nodes beyond V*T unset in the reference: they get huge(1d0) here (no weakening) */
static double twf_mu1(const fault_t *f, int k, double time, double d) {
  const double mus = f->tw[0], mud = f->tw[1], mu0 = f->tw[2], X = f->tw[3], Z = f->tw[4];
  const double V = f->tw[5], L = f->tw[6], T = f->tw[7], Dc = f->tw[8];
  const double big = DBL_MAX;
  const double dx = f->xz[k] - X, dz = f->xz[f->np + k] - Z;
  const double dist = sqrt(dx * dx + dz * dz);
  if (f->twf == 1 || f->twf == 2) {
    double t, r;
    if (f->twf == 1) {
      t = time + (mus - mu0) * L / ((mus - mud) * V);
      if (t > T) t = 0.0;
      r = V * t;
    } else {
      t = time + 0.5 * T * (1.0 - sqrt(1.0 - 4.0 * (mus - mu0) * L / ((mus - mud) * T * V)));
      t = fmin(t, T);
      r = V * t * (1.0 - t / T);
    }
    r = dist - r;
    if (r < -L) return mud;
    if (r <= L) return mus + (mus - mud) / L * r;
    return big;
  }
  if (dist <= V * T && d <= Dc) {
    if (dist < V * time - L) return mud;
    if (dist >= V * time - L && dist <= V * time) return mus + (mus - mud) / L * (dist - V * time);
    return big;
  }
  return big;
}

static double swf_mu1(const fault_t *f, int k) {
  double th = f->theta[k], dc = f->p[0][k], mus = f->p[1][k], mud = f->p[2][k];
  double mu;
  if (f->kind == 1) mu = mus - (mus - mud) * fmin(th / dc, 1.0);
  else if (f->kind == 2) mu = mud - (mud - mus) * exp(-th / dc);
  else mu = mud + (mus - mud) / pow(1.0 + th / dc, f->p[4][k]);
  return mu + f->p[3][k] * th;
}

static double rsf_mu1(const fault_t *f, int k, double v) {
  double dc = f->p[0][k], mus = f->p[1][k], a = f->p[2][k], b = f->p[3][k], vs = f->p[4][k];
  double th = f->theta[k];
  switch (f->kind) {
    case 1: return mus + a * fabs(v) / (fabs(v) + vs) - b * th / (th + dc);
    case 2: case 3: return a * asinh(fabs(v) / (2.0 * vs) * exp((mus + b * log(vs * th / dc)) / a));
    default: return a * asinh(fabs(v) / (2.0 * vs) * exp((mus + b * log(f->p[5][k] * th / dc + 1)) / a));
  }
}

static double rsf_theta1(const fault_t *f, int k, double theta, double v, double dt) {
  double dc = f->p[0][k];
  switch (f->kind) {
    case 1: {
      double tn = theta * f->coeft[k] + f->Tc[k] * fabs(v) * (1.0 - f->coeft[k]);
      return tn < 1.0e-12 ? 0.0 : tn;
    }
    case 2: case 4: {
      double x = fabs(v) / dc;
      double ex = exp(-dt * x);
      if (dt * x > 1e-8) return theta * ex + (1.0 - ex) / x;
      return theta * ex + dt * (1.0 - 0.5 * dt * x);
    }
    default: {
      double tn = dc / fabs(v);
      return tn * pow(theta / tn, exp(-dt / tn));
    }
  }
}

/* nr_fric_func_tau (SRC/bc_dynflt_rsf.f90:533-568) */
static void fric_func(const fault_t *f, int k, double theta, double tau, double tstick,
                      double sigma, double Z, double *fv, double *dfv, double *v) {
  double dc = f->p[0][k], mus = f->p[1][k], a = f->p[2][k], b = f->p[3][k], vs = f->p[4][k];
  double tmp;
  if (f->kind == 4) tmp = mus + b * log(f->p[5][k] * theta / dc + 1.0);
  else tmp = mus + b * log(vs * theta / dc);
  tmp = 2.0 * vs * exp(-tmp / a);
  *v = sinh(tau / (-sigma * a)) * tmp;
  *fv = tstick - Z * (*v) - tau;
  double dv = cosh(tau / (-sigma * a)) * tmp / (-sigma * a);
  *dfv = -Z * dv - 1.0;
}

/* nr_solver (SRC/bc_dynflt_rsf.f90:369-468); returns v of the last evaluation */
static double nr_solver(const fault_t *f, int k, double xL, double xR, double xacc, double theta,
                        double tstick, double sigma, double Z, int *ok) {
  double flow, fhigh, dfdx, v, xl, xr, xlow, xhigh, xest, dx, dxold, fx, tmp;
  *ok = 1;
  fric_func(f, k, theta, xL, tstick, sigma, Z, &flow, &dfdx, &v);
  fric_func(f, k, theta, xR, tstick, sigma, Z, &fhigh, &dfdx, &v);
  xl = xL; xr = xR;
  while (flow * fhigh > 0) {
    xl = xl / 2.0;
    xr = xr * 2.0;
    fric_func(f, k, theta, xl, tstick, sigma, Z, &flow, &dfdx, &v);
    fric_func(f, k, theta, xr, tstick, sigma, Z, &fhigh, &dfdx, &v);
  }
  if (flow == 0) return v;
  if (fhigh == 0) return v;
  if (flow < 0) { xlow = xl; xhigh = xr; } else { xhigh = xl; xlow = xr; }
  xest = 0.5 * (xl + xr);
  dxold = fabs(xr - xl);
  dx = dxold;
  fric_func(f, k, theta, xest, tstick, sigma, Z, &fx, &dfdx, &v);
  for (int is = 1; is <= 200; ++is) {
    if (((xest - xhigh) * dfdx - fx) * ((xest - xlow) * dfdx - fx) > 0 ||
        fabs(2 * fx) > fabs(dxold * dfdx)) {
      dxold = dx;
      dx = 0.5 * (xhigh - xlow);
      xest = xlow + dx;
      if (xlow == xest) return v;
    } else {
      dxold = dx;
      dx = fx / dfdx;
      tmp = xest;
      xest = xest - dx;
      if (tmp == xest) return v;
    }
    if (fabs(dx) < fabs(xacc)) {
      fric_func(f, k, theta, xest, tstick, sigma, Z, &fx, &dfdx, &v);
      return v;
    }
    fric_func(f, k, theta, xest, tstick, sigma, Z, &fx, &dfdx, &v);
    if (fx < 0) xlow = xest; else xhigh = xest;
  }
  *ok = 0;
  return v;
}

/* rsf_update_V (SRC/bc_dynflt_rsf.f90:325-358) for one node */
static double rsf_v1(const fault_t *f, int k, double tstick, double sigma, double theta, double Z,
                     int *ok) {
  double dc = f->p[0][k], mus = f->p[1][k], a = f->p[2][k], b = f->p[3][k], vs = f->p[4][k];
  *ok = 1;
  if (f->kind == 1) {
    /* rsf_mu_no_direct reads the state stored in the friction law (f%theta, the state
       before this step: SRC/bc_dynflt_rsf.f90:189-196,338), not the updated estimate
       passed in; the kind-1 branch does not use its theta argument */
    double th0 = f->theta[k];
    double mnd = mus - b * th0 / (th0 + dc);
    double v = (tstick + sigma * mnd) / Z;
    double tmp = v - vs + sigma * a / Z;
    v = 0.5 * (tmp + sqrt(tmp * tmp + 4.0 * v * vs));
    v = fmax(0.0, v);
    return v < 1.0e-12 ? 0.0 : v;
  }
  double tol = -0.001 * a * sigma;
  double lo = fmin(0.0, tstick), hi = fmax(0.0, tstick);
  return nr_solver(f, k, lo, hi, tol, theta, tstick, sigma, Z, ok);
}

/* get_jump (SRC/bc_dynflt.f90:692-704): x(node2) - x(node1), or -2 x(node1) on a
   one-sided fault */
static double jump(const fault_t *f, const double *x, int k) {
  return f->n2[k] >= 0 ? x[f->n2[k]] - x[f->n1[k]] : -2.0 * x[f->n1[k]];
}

/* BC_DYNFLT_potency (SRC/bc_dynflt.f90:832-855) of the field x (jumps as get_jump):
   P-SV p11, p22, p12; SH p13, p23 (sequential sums, as Fortran's SUM) */
static void potency(const oracle_solver *h, const double *x, double *p) {
  const fault_t *f = h->flt;
  const int np = f->np, N = h->npoin;
  const double *n1x = f->nrm, *n1z = f->nrm + np, *B = f->B;
  if (h->ndof == 2) {
    double s1 = 0, s2 = 0, s3 = 0;
    for (int k = 0; k < np; ++k) s1 += n1x[k] * jump(f, x, k) * B[k];
    for (int k = 0; k < np; ++k) s2 += n1z[k] * jump(f, x + N, k) * B[k];
    for (int k = 0; k < np; ++k) s3 += (n1x[k] * jump(f, x + N, k) + n1z[k] * jump(f, x, k)) * B[k];
    p[0] = s1; p[1] = s2; p[2] = 0.5 * s3;
  } else {
    double s1 = 0, s2 = 0;
    for (int k = 0; k < np; ++k) s1 += n1x[k] * jump(f, x, k) * B[k];
    for (int k = 0; k < np; ++k) s2 += n1z[k] * jump(f, x, k) * B[k];
    p[0] = 0.5 * s1; p[1] = 0.5 * s2;
  }
}

/* one potency row per call (BC_DYNFLT_write, SRC/bc_dynflt.f90:757) */
static void potency_row(oracle_solver *h) {
  fault_t *f = h->flt;
  const int nc = 2 * (h->ndof + 1);
  if (f->pot_rows == f->pot_cap) {
    f->pot_cap = f->pot_cap ? 2 * f->pot_cap : 1024;
    f->pot = (double *)realloc(f->pot, sizeof(double) * nc * f->pot_cap);
  }
  double *r = f->pot + (size_t)f->pot_rows * nc;
  potency(h, h->d, r);
  potency(h, h->v, r + h->ndof + 1);
  f->pot_rows++;
}

/* export_side of get_side (SRC/bc_dynflt.f90:769-812) for the output nodes: rows
   d of side 1, d of side 2, v of side 1, v of side 2, each in the fault frame (t, n)
   (rotate, SRC/bc_dynflt.f90:722-739) for P-SV; side 2 of a fault with one side is
   -(side 1) */
static void record_sides(const oracle_solver *h, double *q) {
  const fault_t *f = h->flt;
  const double *fld[2] = {h->d, h->v};
  int row = 0;
  for (int w = 0; w < 2; ++w)
    for (int side = 1; side <= 2; ++side) {
      int o = 0;
      for (int k = f->oix1 - 1; k <= f->oixn - 1; k += f->oixd, ++o) {
        double x[2] = {0.0, 0.0};
        for (int c = 0; c < f->ndof; ++c) {
          const double *u = fld[w] + (size_t)c * h->npoin;
          x[c] = side == 1 ? u[f->n1[k]] : (f->n2[k] >= 0 ? u[f->n2[k]] : -u[f->n1[k]]);
        }
        if (f->ndof == 1) {
          q[(size_t)row * f->nout + o] = x[0];
        } else {
          const double n1x = f->nrm[k], n1z = f->nrm[f->np + k];
          q[(size_t)row * f->nout + o] = n1z * x[0] - n1x * x[1];
          q[(size_t)(row + 1) * f->nout + o] = n1x * x[0] + n1z * x[1];
        }
      }
      row += f->ndof;
    }
}

static void fault_record(oracle_solver *h) {
  fault_t *f = h->flt;
  potency_row(h);
  if (h->it < f->oit) return;
  if (f->taken < f->nsamp) {
    double *r = f->rec + (size_t)f->taken * 6 * f->nout;
    int o = 0;
    for (int k = f->oix1 - 1; k <= f->oixn - 1; k += f->oixd, ++o) {
      r[0 * f->nout + o] = f->D[k];
      r[1 * f->nout + o] = f->V[k];
      r[2 * f->nout + o] = f->T[k];
      r[3 * f->nout + o] = f->T[f->np + k];
      r[4 * f->nout + o] = f->MU[k];
      r[5 * f->nout + o] = f->Tstick[k];
    }
    if (f->osides) record_sides(h, f->srec + (size_t)f->taken * 4 * f->ndof * f->nout);
    f->taken++;
  }
  f->oit += f->oitd;
}

/* superimposed (law 1, 2) or sole (law 3) time-weakening friction (BC_DYNFLT_TWF,
   SRC/bc_dynflt_twf.f90); par = MuS, MuD, Mu0, X, Z, V, L, T, Dc; coord [2][np] */
int oracle_set_fault_twf(oracle_solver *h, int kind, const double *par, const double *coord) {
  fault_t *f = h->flt;
  if (!f) return fail(h, 4, "set_fault_twf: no fault");
  if (kind < 1 || kind > 3 || !par || !coord) return fail(h, 3, "set_fault_twf: bad arguments");
  f->twf = kind;
  memcpy(f->tw, par, sizeof f->tw);
  free(f->xz);
  f->xz = dup_d(coord, 2 * (size_t)f->np);
  /* initial friction (SRC/bc_dynflt.f90:404-413) */
  for (int k = 0; k < f->np; ++k) {
    const double mt = twf_mu1(f, k, 0.0, f->D[k]);
    f->MU[k] = f->law == 3 ? mt : fmin(f->MU[k], mt);
  }
  if (f->taken) { /* the friction row of the record of step 0, taken by set_fault */
    int o = 0;
    for (int k = f->oix1 - 1; k <= f->oixn - 1; k += f->oixd, ++o) f->rec[4 * f->nout + o] = f->MU[k];
  }
  return 0;
}

/* BC_DYNFLT osides = T: every fault record also keeps the displacement and velocity of
   each side in the fault frame (SRC/bc_dynflt.f90:769-812) */
int oracle_set_fault_normal(oracle_solver *h, int kind, double coef, double v) {
  if (!h->flt) return fail(h, 4, "set_fault_normal: call set_fault first");
  if (h->it != 0) return fail(h, 4, "set_fault_normal: the solver has already been stepped");
  if (kind < 0 || kind > 3) return fail(h, 3, "set_fault_normal: kind must be 0 to 3");
  h->flt->nor_kind = kind;
  h->flt->nor_coef = coef;
  h->flt->nor_v = v;
  return 0;
}

int oracle_set_periodic(oracle_solver *h, int n, const int *master, const int *slave) {
  if (n <= 0 || !master || !slave) return fail(h, 3, "set_periodic: bad arguments");
  if (h->nper) return fail(h, 3, "set_periodic: one periodic boundary per solver");
  char *mark = (char *)calloc((size_t)h->npoin, 1);
  for (int k = 0; k < n; ++k) {
    int m = master[k] - 1, s = slave[k] - 1;
    if (m < 0 || m >= h->npoin || s < 0 || s >= h->npoin || mark[m] || mark[s] || m == s) {
      free(mark);
      return fail(h, 3, "set_periodic: nodes out of range or repeated");
    }
    mark[m] = mark[s] = 1;
  }
  free(mark);
  h->per_m = (int *)malloc((size_t)n * sizeof(int));
  h->per_s = (int *)malloc((size_t)n * sizeof(int));
  for (int k = 0; k < n; ++k) {
    h->per_m[k] = master[k] - 1;
    h->per_s[k] = slave[k] - 1;
  }
  h->nper = n;
  return 0;
}

int oracle_set_fault_sides(oracle_solver *h, int on) {
  fault_t *f = h->flt;
  if (!f) return fail(h, 4, "set_fault_sides: no fault");
  if (f->taken > 1 || h->it) return fail(h, 4, "set_fault_sides: call it before the first step");
  f->osides = on != 0;
  free(f->srec);
  f->srec = on ? dup_d(NULL, (size_t)f->nsamp * 4 * f->ndof * f->nout) : NULL;
  if (on && f->taken) record_sides(h, f->srec); /* the record of step 0, taken by set_fault */
  return 0;
}

int oracle_fetch_fault_sides(oracle_solver *h, double *out, int *taken) {
  fault_t *f = h->flt;
  if (!f || !f->osides) return fail(h, 4, "fetch_fault_sides: no per-side records");
  memcpy(out, f->srec, sizeof(double) * f->taken * 4 * f->ndof * f->nout);
  if (taken) *taken = f->taken;
  return 0;
}

int oracle_set_fault(oracle_solver *h, const sem2d_fault_desc *d, int nsamp) {
  fault_t *f = (fault_t *)calloc(1, sizeof *f);
  f->nor_kind = 1;
  int np = d->npoin, nd = h->ndof;
  f->np = np; f->ndof = nd;
  f->n1 = (int *)malloc(sizeof(int) * np);
  f->n2 = (int *)malloc(sizeof(int) * np);
  /* a fault on a domain boundary (one side, symmetry: SRC/bc_dynflt.f90:256-259) has no
     node2 and no invM2: n2 = -1 */
  for (int k = 0; k < np; ++k) { f->n1[k] = d->node1[k] - 1; f->n2[k] = d->node2 ? d->node2[k] - 1 : -1; }
  f->nrm = dup_d(d->n1, 2 * (size_t)np);
  f->B = dup_d(d->B, (size_t)nd * np);
  f->invM1 = dup_d(d->invM1, (size_t)nd * np);
  f->invM2 = d->node2 ? dup_d(d->invM2, (size_t)nd * np) : NULL;
  f->Z = dup_d(d->Z, (size_t)nd * np);
  f->T0 = dup_d(d->T0, 2 * (size_t)np);
  f->coh = dup_d(d->cohesion, np);
  f->a2v = d->coefA2V; f->a2d = d->coefA2D;
  f->opening = d->allow_opening; f->law = d->law; f->kind = d->kind; f->healing = d->healing;
  for (int q = 0; q < 6; ++q) f->p[q] = d->p[q] ? dup_d(d->p[q], np) : dup_d(NULL, np);
  f->theta = dup_d(d->theta0, np);
  f->Tc = dup_d(NULL, np); f->coeft = dup_d(NULL, np);
  if (f->law == 2)
    for (int k = 0; k < np; ++k) {
      f->Tc[k] = f->p[0][k] / f->p[4][k];
      f->coeft[k] = exp(-h->dt / f->Tc[k]);
    }
  f->D = dup_d(NULL, (size_t)nd * np);
  f->V = dup_d(NULL, (size_t)nd * np);
  if (d->V0) memcpy(f->V, d->V0, sizeof(double) * np);
  f->T = dup_d(NULL, 2 * (size_t)np);
  f->Tstick = dup_d(NULL, 2 * (size_t)np);
  f->MU = dup_d(NULL, np);
  f->sigma = dup_d(NULL, np);
  for (int k = 0; k < np; ++k) {
    f->MU[k] = f->law == 1 ? swf_mu1(f, k) : (f->law == 2 ? rsf_mu1(f, k, f->V[k]) : DBL_MAX);
    f->sigma[k] = f->T0[np + k];
  }
  f->oit = d->oit; f->oitd = d->oitd; f->oix1 = d->oix1; f->oixn = d->oixn; f->oixd = d->oixd;
  f->nout = (f->oixn - f->oix1) / f->oixd + 1;
  f->nsamp = nsamp;
  f->rec = dup_d(NULL, (size_t)nsamp * 6 * f->nout);
  h->flt = f;
  fault_record(h);
  return 0;
}

/* ---------------- element forces ---------------- */
static void mm(int ng, const double A[NGMAX][NGMAX], const double B[NGMAX][NGMAX], double C[NGMAX][NGMAX]) {
  /* My_MATMUL / mxm: C(i,j) = sum_k A(i,k) B(k,j) in k order; arrays indexed [col][row] =
     Fortran (row,col) */
  for (int j = 0; j < ng; ++j)
    for (int i = 0; i < ng; ++i) {
      double c = 0.0;
      for (int k = 0; k < ng; ++k) c = c + A[k][i] * B[j][k];
      C[j][i] = c;
    }
}

/* Element forces of one element.  NGLL = 5 (OPT_NGLL): ELAST_KD2_PSV / ELAST_KD2_SH
   (SRC/mat_elastic.f90:589-641,739-775); other NGLL: ELAST_KD1_PSV / ELAST_KD1_SH
   (SRC/mat_elastic.f90:464-518,645-681), which differ from the KD2 forms only in the
   second term of the x force of flat P-SV elements: a4 dUx/deta + a5 dUz/dxi instead of
   a4 (dUx/deta + dUz/dxi).  Element arrays are stored [j][i] (Fortran (i,j)). */
static void kd(const oracle_solver *h, const double *a, double U[2][NGMAX][NGMAX], double F[2][NGMAX][NGMAX]) {
  const int ng = h->ng;
  const int kd1 = ng != 5;
  double Hm[NGMAX][NGMAX], Ht[NGMAX][NGMAX]; /* [col][row] */
  for (int i = 0; i < ng; ++i)
    for (int j = 0; j < ng; ++j) { Hm[j][i] = h->H[i][j]; Ht[j][i] = h->H[j][i]; }
#define A(k, i, j) a[((k) * ng + (j)) * ng + (i)]
#define EACH for (int j = 0; j < ng; ++j) for (int i = 0; i < ng; ++i)
  double tmp[NGMAX][NGMAX], t2[NGMAX][NGMAX];
  if (h->ndof == 2) {
    double dxx[NGMAX][NGMAX], dzx[NGMAX][NGMAX], dxe[NGMAX][NGMAX], dze[NGMAX][NGMAX];
    mm(ng, Ht, U[0], dxx);
    mm(ng, Ht, U[1], dzx);
    mm(ng, U[0], Hm, dxe);
    mm(ng, U[1], Hm, dze);
    if (h->nelast == 6) {
      EACH tmp[j][i] = A(0, i, j) * dxx[j][i] + A(1, i, j) * dze[j][i];
      mm(ng, Hm, tmp, F[0]);
      if (kd1) {
        EACH tmp[j][i] = A(3, i, j) * dxe[j][i] + A(4, i, j) * dzx[j][i];
      } else {
        EACH tmp[j][i] = A(3, i, j) * (dxe[j][i] + dzx[j][i]);
      }
      mm(ng, tmp, Ht, t2);
      EACH F[0][j][i] = F[0][j][i] + t2[j][i];
      EACH tmp[j][i] = A(4, i, j) * dxe[j][i] + A(5, i, j) * dzx[j][i];
      mm(ng, Hm, tmp, F[1]);
      EACH tmp[j][i] = A(1, i, j) * dxx[j][i] + A(2, i, j) * dze[j][i];
      mm(ng, tmp, Ht, t2);
      EACH F[1][j][i] = F[1][j][i] + t2[j][i];
    } else {
      EACH tmp[j][i] = A(0, i, j) * dxx[j][i] + A(6, i, j) * dxe[j][i] + A(7, i, j) * dzx[j][i] +
                       A(1, i, j) * dze[j][i];
      mm(ng, Hm, tmp, F[0]);
      EACH tmp[j][i] = A(6, i, j) * dxx[j][i] + A(3, i, j) * dxe[j][i] + A(4, i, j) * dzx[j][i] +
                       A(8, i, j) * dze[j][i];
      mm(ng, tmp, Ht, t2);
      EACH F[0][j][i] = F[0][j][i] + t2[j][i];
      EACH tmp[j][i] = A(7, i, j) * dxx[j][i] + A(4, i, j) * dxe[j][i] + A(5, i, j) * dzx[j][i] +
                       A(9, i, j) * dze[j][i];
      mm(ng, Hm, tmp, F[1]);
      EACH tmp[j][i] = A(1, i, j) * dxx[j][i] + A(8, i, j) * dxe[j][i] + A(9, i, j) * dzx[j][i] +
                       A(2, i, j) * dze[j][i];
      mm(ng, tmp, Ht, t2);
      EACH F[1][j][i] = F[1][j][i] + t2[j][i];
    }
  } else {
    double dx[NGMAX][NGMAX], de[NGMAX][NGMAX];
    mm(ng, Ht, U[0], dx);
    mm(ng, U[0], Hm, de);
    if (h->nelast == 2) {
      EACH tmp[j][i] = A(0, i, j) * dx[j][i];
      mm(ng, Hm, tmp, F[0]);
      EACH tmp[j][i] = A(1, i, j) * de[j][i];
    } else {
      EACH tmp[j][i] = A(0, i, j) * dx[j][i] + A(2, i, j) * de[j][i];
      mm(ng, Hm, tmp, F[0]);
      EACH tmp[j][i] = A(2, i, j) * dx[j][i] + A(1, i, j) * de[j][i];
    }
    mm(ng, tmp, Ht, t2);
    EACH F[0][j][i] = F[0][j][i] + t2[j][i];
  }
#undef EACH
#undef A
}

/* element visiting order of the assembly: the reference's (element order), or one of
   the perturbed orders described at the top of this file */
static void make_order(oracle_solver *h) {
  const int ne = h->nelem;
  h->eorder = (int *)malloc(sizeof(int) * (size_t)ne);
  if (ORACLE_ORDER == 1) {
    for (int e = 0; e < ne; ++e) h->eorder[e] = ne - 1 - e;
  } else if (ORACLE_ORDER == 2) {
    unsigned long long *used = (unsigned long long *)calloc((size_t)h->npoin, sizeof *used);
    unsigned char *col = (unsigned char *)malloc((size_t)ne);
    int ncol = 0;
    for (int e = 0; e < ne; ++e) {
      unsigned long long m = 0;
      const int *ib = h->ibool + (size_t)e * h->npt;
      for (int k = 0; k < h->npt; ++k) m |= used[ib[k]];
      int c = 0;
      while (m >> c & 1ull) ++c;
      col[e] = (unsigned char)c;
      if (c + 1 > ncol) ncol = c + 1;
      for (int k = 0; k < h->npt; ++k) used[ib[k]] |= 1ull << c;
    }
    int q = 0;
    for (int c = 0; c < ncol; ++c)
      for (int e = 0; e < ne; ++e)
        if (col[e] == c) h->eorder[q++] = e;
    free(used);
    free(col);
  } else {
    for (int e = 0; e < ne; ++e) h->eorder[e] = e;
  }
}

/* compute_Fint (SRC/solver.f90:273-320); v != NULL adds the Kelvin-Voigt term
   d + eta v in the viscous elements (MAT_KV_add_etav, SRC/mat_kelvin_voigt.f90:139-150).
   The element forces are independent of each other and are computed in parallel into
   a per-element buffer; the assembly f(ibool) = f(ibool) + floc then runs sequentially
   in the element order, so the result is the reference's whatever the thread count. */
static void fint(oracle_solver *h, const double *d, double *f, const double *v) {
  size_t np = h->npoin;
  const int nd = h->ndof, ne = h->nelem;
  if (!h->eorder) make_order(h);
  if (!h->floc) h->floc = (double *)malloc(sizeof(double) * (size_t)ne * nd * h->npt);
  double *floc = h->floc;
#pragma omp parallel for schedule(static) if (ne >= 2048)
  for (int e = 0; e < ne; ++e) {
    double U[2][NGMAX][NGMAX], F[2][NGMAX][NGMAX];
    const int *ib = h->ibool + (size_t)e * h->npt;
    for (int c = 0; c < nd; ++c)
      for (int j = 0; j < h->ng; ++j)
        for (int i = 0; i < h->ng; ++i) U[c][j][i] = d[c * np + ib[j * h->ng + i]];
    if (v && h->kv_eta) {
      const double *eta = h->kv_eta + (size_t)e * h->npt;
      for (int c = 0; c < nd; ++c)
        for (int j = 0; j < h->ng; ++j)
          for (int i = 0; i < h->ng; ++i) U[c][j][i] = U[c][j][i] + eta[j * h->ng + i] * v[c * np + ib[j * h->ng + i]];
    }
    const double *a = h->coefs + (h->shared ? 0 : (size_t)e * h->nelast * h->npt);
    kd(h, a, U, F);
    if (h->beta25) { /* MAT_ELAST_add_25D_f (SRC/mat_elastic.f90:447-459) */
      const double *b = h->beta25 + (size_t)e * h->npt;
      for (int c = 0; c < nd; ++c)
        for (int j = 0; j < h->ng; ++j)
          for (int i = 0; i < h->ng; ++i) F[c][j][i] = F[c][j][i] - b[j * h->ng + i] * U[c][j][i];
    }
    double *fl = floc + (size_t)e * nd * h->npt;
    for (int c = 0; c < nd; ++c)
      for (int j = 0; j < h->ng; ++j)
        for (int i = 0; i < h->ng; ++i) fl[(c * h->ng + j) * h->ng + i] = F[c][j][i];
  }
  memset(f, 0, sizeof(double) * np * nd);
  for (int q = 0; q < ne; ++q) {
    const int e = h->eorder[q];
    const int *ib = h->ibool + (size_t)e * h->npt;
    const double *fl = floc + (size_t)e * nd * h->npt;
    for (int j = 0; j < h->ng; ++j)
      for (int i = 0; i < h->ng; ++i) {
        int k = ib[j * h->ng + i];
        for (int c = 0; c < nd; ++c) f[c * np + k] = f[c * np + k] + fl[(c * h->ng + j) * h->ng + i];
      }
  }
}

int oracle_compute_fint(oracle_solver *h, const double *d, double *f) {
  fint(h, d, f, NULL);
  return 0;
}

int oracle_set_kelvin_voigt(oracle_solver *h, const double *eta) {
  h->kv_eta = dup_d(eta, (size_t)h->nelem * h->npt);
  return 0;
}

int oracle_set_25d(oracle_solver *h, const double *beta) {
  if (!beta) return fail(h, 3, "set_25d: null beta");
  h->beta25 = dup_d(beta, (size_t)h->nelem * h->npt);
  return 0;
}

static void fault_apply(oracle_solver *h, double *MxA, double time) {
  fault_t *f = h->flt;
  int np = f->np, nd = h->ndof;
  size_t N = h->npoin;
  for (int k = 0; k < np; ++k) {
    int i1 = f->n1[k], i2 = f->n2[k];
    double dD[2] = {0, 0}, dV[2] = {0, 0}, dA[2] = {0, 0}, T[2] = {0, 0}, Ts[2];
    for (int c = 0; c < nd; ++c) {
      if (i2 >= 0) {
        dD[c] = h->d[c * N + i2] - h->d[c * N + i1];
        dV[c] = h->v[c * N + i2] - h->v[c * N + i1];
        dA[c] = f->invM2[c * np + k] * MxA[c * N + i2] - f->invM1[c * np + k] * MxA[c * N + i1];
      } else { /* one side: get_jump, get_weighted_jump (SRC/bc_dynflt.f90:692-719) */
        dD[c] = -2.0 * h->d[c * N + i1];
        dV[c] = -2.0 * h->v[c * N + i1];
        dA[c] = -2.0 * f->invM1[c * np + k] * MxA[c * N + i1];
      }
      T[c] = f->Z[c * np + k] * (dV[c] + f->a2v * dA[c]);
    }
    double nx = f->nrm[k], nz = f->nrm[np + k];
    if (nd == 2) {
      double r0, r1;
      r0 = nz * dD[0] - nx * dD[1]; r1 = nx * dD[0] + nz * dD[1]; dD[0] = r0; dD[1] = r1;
      r0 = nz * dV[0] - nx * dV[1]; r1 = nx * dV[0] + nz * dV[1]; dV[0] = r0; dV[1] = r1;
      r0 = nz * dA[0] - nx * dA[1]; r1 = nx * dA[0] + nz * dA[1]; dA[0] = r0; dA[1] = r1;
      r0 = nz * T[0] - nx * T[1]; r1 = nx * T[0] + nz * T[1]; T[0] = r0; T[1] = r1;
    }
    if (nd == 1 || i2 < 0) T[1] = 0.0; /* SRC/bc_dynflt.f90:604 */
    T[0] = T[0] + f->T0[k];
    T[1] = T[1] + f->T0[np + k];
    if (f->opening) T[1] = fmin(T[1], 0.0);
    /* normal_update (SRC/bc_dynflt_normal.f90:122-140), dV(:,1) the trial slip rate */
    if (f->nor_kind == 1) f->sigma[k] = T[1];
    else if (f->nor_kind == 2) f->sigma[k] = T[1] + f->nor_coef * (f->sigma[k] - T[1]);
    else if (f->nor_kind == 3)
      f->sigma[k] = T[1] + exp(-(fabs(dV[0]) + f->nor_v) * f->nor_coef) * (f->sigma[k] - T[1]);
    double sigma = f->sigma[k];
    Ts[0] = T[0]; Ts[1] = T[1];
    if (f->law == 2) {
      double vold = f->V[k], Z1 = f->Z[k];
      int ok;
      double th = rsf_theta1(f, k, f->theta[k], vold, h->dt);
      double vn = rsf_v1(f, k, T[0], sigma, th, Z1, &ok);
      th = rsf_theta1(f, k, f->theta[k], 0.5 * (vold + vn), h->dt);
      vn = rsf_v1(f, k, T[0], sigma, th, Z1, &ok);
      f->theta[k] = th;
      f->V[k] = vn;
      f->MU[k] = rsf_mu1(f, k, f->V[k]);
      /* superimposed time weakening (SRC/bc_dynflt.f90:632) */
      if (f->twf) f->MU[k] = fmin(f->MU[k], twf_mu1(f, k, time, f->D[k]));
      double strength = -f->MU[k] * sigma;
      T[0] = copysign(fabs(strength), T[0]);
    } else if (f->law == 3) { /* time weakening alone (SRC/bc_dynflt.f90:656-657) */
      f->MU[k] = twf_mu1(f, k, time, f->D[k]);
      double strength = f->coh[k] - f->MU[k] * sigma;
      T[0] = copysign(fmin(fabs(T[0]), strength), T[0]);
    } else {
      if (f->a2d == 0.0) {
        if (f->healing) {
          f->theta[k] = f->theta[k] + fabs(dV[0]) * h->dt;
          if (fabs(dV[0]) < 1e-14) f->theta[k] = 0.0;
        } else {
          f->theta[k] = fabs(dD[0]);
        }
      } else {
        f->theta[k] = fabs(f->D[k]);
      }
      f->MU[k] = swf_mu1(f, k);
      if (f->twf) f->MU[k] = fmin(f->MU[k], twf_mu1(f, k, time, f->D[k])); /* :653 */
      double strength = f->coh[k] - f->MU[k] * sigma;
      T[0] = copysign(fmin(fabs(T[0]), strength), T[0]);
    }
    T[0] = T[0] - f->T0[k];
    T[1] = T[1] - f->T0[np + k];
    Ts[0] = Ts[0] - f->T0[k];
    Ts[1] = Ts[1] - f->T0[np + k];
    f->T[k] = T[0]; f->T[np + k] = T[1];
    f->Tstick[k] = Ts[0]; f->Tstick[np + k] = Ts[1];
    double Tx[2] = {T[0], T[1]};
    if (nd == 2) {
      Tx[0] = nz * T[0] + nx * T[1];
      Tx[1] = -nx * T[0] + nz * T[1];
    }
    for (int c = 0; c < nd; ++c) {
      MxA[c * N + i1] = MxA[c * N + i1] + f->B[c * np + k] * Tx[c];
      if (i2 >= 0) MxA[c * N + i2] = MxA[c * N + i2] - f->B[c * np + k] * Tx[c];
    }
    for (int c = 0; c < nd; ++c) {
      double tf = (c == 0) ? f->T[k] : f->T[np + k];
      dA[c] = dA[c] - tf / (f->Z[c * np + k] * f->a2v);
      f->D[c * np + k] = dD[c] + f->a2d * dA[c];
      f->V[c * np + k] = dV[c] + f->a2v * dA[c];
    }
  }
}

static void one_step(oracle_solver *h) {
  size_t n = (size_t)h->ndof * h->npoin, N = h->npoin;
  double dt = h->dt, beta = h->beta, gam = h->gamma;
  double *d = h->d, *v = h->v, *a = h->a;
  double c1 = (0.5 - beta) * dt * dt, c2 = (1.0 - gam) * dt;
  double *df = d, *vf = v; /* the fields the forces see (displ_alpha, veloc_alpha) */
  if (h->leapfrog) {
    /* solve_leapfrog (SRC/solver.f90:151): d = d + dt*v_mid */
#pragma omp parallel for schedule(static) if (n >= 65536)
    for (size_t k = 0; k < n; ++k) d[k] = d[k] + dt * v[k];
  } else if (h->hht) {
    /* solve_HHT_alpha (SRC/solver.f90:108-113) */
    double al = h->alpha;
    df = h->da;
    vf = h->va;
    for (size_t k = 0; k < n; ++k) {
      df[k] = d[k];
      vf[k] = v[k];
      d[k] = d[k] + dt * v[k] + c1 * a[k];
      v[k] = v[k] + c2 * a[k];
      df[k] = al * d[k] + (1.0 - al) * df[k];
      vf[k] = al * v[k] + (1.0 - al) * vf[k];
    }
  } else {
    /* solve_Newmark predictors (SRC/solver.f90:58-60) */
#pragma omp parallel for schedule(static) if (n >= 65536)
    for (size_t k = 0; k < n; ++k) {
      d[k] = d[k] + dt * v[k] + c1 * a[k];
      v[k] = v[k] + c2 * a[k];
    }
  }
  double *f = a; /* f => a (SRC/solver.f90:52) */
  fint(h, df, f, vf);
  int it = h->it + 1;
  for (int s = 0; s < h->nsrc; ++s) {
    source_t *S = &h->src[s];
    double amp = (it - 1 < S->nt) ? S->ampl[it - 1] : 0.0;
    if (S->kind == 1) {
      if (h->ndof == 1) f[S->iglob] = f[S->iglob] + amp;
      else {
        f[S->iglob] = f[S->iglob] + S->dir[0] * amp;
        f[N + S->iglob] = f[N + S->iglob] + S->dir[1] * amp;
      }
    } else {
      for (int k = 0; k < S->nel; ++k) {
        for (int c = 0; c < h->ndof; ++c)
          for (int q = 0; q < h->ng; ++q) {
            int g = S->ixi[k * h->ng + q];
            f[c * N + g] = f[c * N + g] + amp * S->cxi[((size_t)k * h->ndof + c) * h->ng + q];
          }
        for (int c = 0; c < h->ndof; ++c)
          for (int q = 0; q < h->ng; ++q) {
            int g = S->ieta[k * h->ng + q];
            f[c * N + g] = f[c * N + g] + amp * S->ceta[((size_t)k * h->ndof + c) * h->ng + q];
          }
      }
    }
  }
  /* periodic boundaries first (SRC/bc_gen.f90:273-275): assemble in the master nodes,
     then copy to the slave nodes (BC_PERIO_set_field, SRC/bc_periodic.f90:78-90) */
  for (int c = 0; c < h->ndof; ++c) {
    for (int k = 0; k < h->nper; ++k) f[c * N + h->per_m[k]] = f[c * N + h->per_m[k]] + f[c * N + h->per_s[k]];
    for (int k = 0; k < h->nper; ++k) f[c * N + h->per_s[k]] = f[c * N + h->per_m[k]];
  }
  double tbc = (double)it * dt; /* time = it*dt (SRC/main.f90:53) */
  if (h->hht) tbc = tbc + (h->alpha - 1.0) * dt; /* t_alpha (SRC/solver.f90:114-119) */
  for (int b = 0; b < h->nabso; ++b) {
    abso_t *B = &h->abso[b];
    if (B->wxz) {
      /* incident wave (SRC/bc_abso.f90:300-306, SRC/src_gen.f90:322-339): MxA - C (V - Vin) + B Tin */
      const double *w = B->wpar;
      for (int k = 0; k < B->n; ++k) {
        double phase = tbc - (B->wxz[k] - w[2]) * w[0] - (B->wxz[B->n + k] - w[3]) * w[1];
        phase = phase - w[4];
        double arg = 3.141592653589793 * w[5] * (phase - w[6]);
        arg = arg * arg;
        double amp = -w[7] * (1.0 - 2.0 * arg) * exp(-arg);
        for (int c = 0; c < h->ndof; ++c) {
          int g = B->node[k];
          double vin = amp * B->wV[c * B->n + k], tin = amp * B->wT[c * B->n + k];
          f[c * N + g] = f[c * N + g] - B->C[c * B->n + k] * (vf[c * N + g] - vin) + B->wB[k] * tin;
        }
      }
    } else if (B->nrm) {
      /* not flat (SRC/bc_abso.f90:313-318): Vn = (V.n) n, MxA -= C1 Vn + C2 (V - Vn) */
      for (int k = 0; k < B->n; ++k) {
        int g = B->node[k];
        double n1 = B->nrm[k], n2 = B->nrm[B->n + k];
        double dot = vf[g] * n1 + vf[N + g] * n2;
        double vn1 = dot * n1, vn2 = dot * n2;
        f[g] = f[g] - B->C[k] * vn1 - B->C[B->n + k] * (vf[g] - vn1);
        f[N + g] = f[N + g] - B->C[k] * vn2 - B->C[B->n + k] * (vf[N + g] - vn2);
      }
    } else {
      for (int c = 0; c < h->ndof; ++c)
        for (int k = 0; k < B->n; ++k) {
          int g = B->node[k];
          f[c * N + g] = f[c * N + g] - B->C[c * B->n + k] * vf[c * N + g];
        }
    }
    if (B->nbe) {
      /* Stacey (SRC/bc_abso.f90:321-333): KxD(k(:),:) += K(:,:,e) * matmul(Ht, D(nodes(k),:))
         over the boundary elements in order, then MxA(nodes,:) -= KxD; D is the predicted
         displacement, Ht(k,j) = h'_j(x_k) */
      double *kxd = (double *)calloc((size_t)B->n * 2, sizeof(double));
      for (int e = 0; e < B->nbe; ++e) {
        const int *en = &B->enode[e * h->ng];
        for (int c = 0; c < 2; ++c)
          for (int kk = 0; kk < h->ng; ++kk) {
            double s = 0.0;
            for (int j = 0; j < h->ng; ++j) s = s + h->H[j][kk] * df[c * N + B->node[en[j]]];
            kxd[en[kk] * 2 + c] = kxd[en[kk] * 2 + c] + B->K[((size_t)e * h->ng + kk) * 2 + c] * s;
          }
      }
      if (B->perends) { /* SRC/bc_abso.f90:325-328 */
        for (int c = 0; c < 2; ++c) {
          kxd[c] = kxd[c] + kxd[(B->n - 1) * 2 + c];
          kxd[(B->n - 1) * 2 + c] = kxd[c];
        }
      }
      for (int c = 0; c < 2; ++c)
        for (int k = 0; k < B->n; ++k) {
          int g = B->node[k];
          f[c * N + g] = f[c * N + g] - kxd[k * 2 + c];
        }
      free(kxd);
    }
  }
  /* the other boundary conditions in their order (SRC/bc_gen.f90:273-281) */
  for (int b = 0; b < h->ndir; ++b)
    if (!h->dir_after_fault[b]) dirichlet_apply(h, b, f);
  if (h->nlsf) {
    /* BC_LSF_apply (SRC/bc_lsf.f90:161-186), D the displacement the forces see */
    int n = h->nlsf, nd = h->ndof;
    double *T = (double *)calloc((size_t)2 * n, sizeof(double));
    for (int k = 0; k < n; ++k) {
      int i1 = h->lsf1[k], i2 = h->lsf2[k];
      double t[2] = {0, 0}, dd[2] = {0, 0};
      for (int c = 0; c < nd; ++c) {
        t[c] = h->lsf_W2[c * n + k] * f[c * N + i2] - h->lsf_W1[c * n + k] * f[c * N + i1];
        dd[c] = df[c * N + i2] - df[c * N + i1];
      }
      if (nd == 2) {
        double nx = h->lsf_nrm[k], nz = h->lsf_nrm[n + k];
        double t0 = nz * t[0] - nx * t[1], t1 = nx * t[0] + nz * t[1];
        double d0 = nz * dd[0] - nx * dd[1], d1 = nx * dd[0] + nz * dd[1];
        if (h->lsf_kt < DBL_MAX) t0 = h->lsf_kt * d0;
        if (h->lsf_kn < DBL_MAX) t1 = h->lsf_kn * d1;
        t[0] = nz * t0 + nx * t1;
        t[1] = -nx * t0 + nz * t1;
      } else if (h->lsf_kt < DBL_MAX) {
        t[0] = h->lsf_kt * dd[0];
      }
      T[2 * k] = t[0];
      T[2 * k + 1] = t[1];
    }
    for (int c = 0; c < nd; ++c) {
      for (int k = 0; k < n; ++k) f[c * N + h->lsf1[k]] = f[c * N + h->lsf1[k]] + h->lsf_B[c * n + k] * T[2 * k + c];
      for (int k = 0; k < n; ++k) f[c * N + h->lsf2[k]] = f[c * N + h->lsf2[k]] - h->lsf_B[c * n + k] * T[2 * k + c];
    }
    free(T);
  }
  if (h->flt) fault_apply(h, f, tbc);
  for (int b = 0; b < h->ndir; ++b)
    if (h->dir_after_fault[b]) dirichlet_apply(h, b, f);
  double c3 = gam * dt, c4 = beta * dt * dt;
  if (h->leapfrog) {
    /* a = rmass*f; v_mid = v_mid + dt*a (SRC/solver.f90:157-158) */
#pragma omp parallel for schedule(static) if (n >= 65536)
    for (size_t k = 0; k < n; ++k) {
      a[k] = h->rmass[k] * f[k];
      v[k] = v[k] + dt * a[k];
    }
  } else {
    /* a = f*rmass; correctors (SRC/solver.f90:77-82) */
#pragma omp parallel for schedule(static) if (n >= 65536)
    for (size_t k = 0; k < n; ++k) {
      a[k] = f[k] * h->rmass[k];
      v[k] = v[k] + c3 * a[k];
      d[k] = d[k] + c4 * a[k];
    }
  }
  h->it = it;
  rec_store(h);
  if (h->flt) fault_record(h);
}

/* maxval(abs(veloc)), maxval(abs(displ)) (SRC/main.f90:45,73-76) and the number of
   values of d, v, a that are not finite */
static void field_info(const oracle_solver *h, double *vmax, double *dmax, int *bad) {
  size_t n = (size_t)h->ndof * h->npoin;
  double vm = 0.0, dm = 0.0;
  int nb = 0;
  for (size_t k = 0; k < n; ++k) {
    nb += !isfinite(h->d[k]) + !isfinite(h->v[k]) + !isfinite(h->a[k]);
    if (fabs(h->v[k]) > vm) vm = fabs(h->v[k]);
    if (fabs(h->d[k]) > dm) dm = fabs(h->d[k]);
  }
  *vmax = vm; *dmax = dm; *bad = nb;
}

int oracle_step(oracle_solver *h, int nsteps) {
  for (int s = 0; s < nsteps; ++s) {
    one_step(h);
    if (h->itinfo > 0 && h->it % h->itinfo == 0) {
      if (h->info_rows == h->info_cap) {
        h->info_cap = h->info_cap ? 2 * h->info_cap : 64;
        h->info = (double *)realloc(h->info, sizeof(double) * 4 * h->info_cap);
      }
      double *r = &h->info[4 * h->info_rows++];
      int bad;
      r[0] = h->it; r[1] = h->it * h->dt;
      field_info(h, &r[2], &r[3], &bad);
      if (bad && !h->first_bad) h->first_bad = h->it;
    }
    if (h->first_bad) {
      snprintf(h->err, sizeof h->err, "non-finite field values at step %d", h->first_bad);
      return 21;
    }
  }
  return 0;
}

/* kind 0: Newmark (solve_Newmark, SRC/solver.f90:42-84), 1: leapfrog (solve_leapfrog,
   SRC/solver.f90:140-160); the dispatch of solve (SRC/solver.f90:20-37) */
int oracle_set_abso_wave(oracle_solver *h, int n, const double *xz, const double *vdir, const double *tdir,
                         const double *Bw, const double *par) {
  if (h->nabso < 1) return fail(h, 3, "set_abso_wave: no absorbing boundary");
  abso_t *b = &h->abso[h->nabso - 1];
  if (n != b->n || !xz || !vdir || !tdir || !Bw || !par) return fail(h, 3, "set_abso_wave: bad arguments");
  b->wxz = dup_d(xz, (size_t)2 * n);
  b->wV = dup_d(vdir, (size_t)h->ndof * n);
  b->wT = dup_d(tdir, (size_t)h->ndof * n);
  b->wB = dup_d(Bw, (size_t)n);
  for (int q = 0; q < 8; ++q) b->wpar[q] = par[q];
  return 0;
}

int oracle_set_abso_periodic(oracle_solver *h) {
  if (h->nabso < 1 || !h->abso[h->nabso - 1].nbe) return fail(h, 3, "set_abso_periodic: no Stacey boundary");
  h->abso[h->nabso - 1].perends = 1;
  return 0;
}

int oracle_set_lsf(oracle_solver *h, int n, const int *node1, const int *node2, const double *nrm,
                   const double *Bv, const double *W1, const double *W2, double kt, double kn) {
  if (n <= 0 || !node1 || !node2 || !nrm || !Bv || !W1 || !W2) return fail(h, 3, "set_lsf: bad arguments");
  if (h->nlsf) return fail(h, 3, "set_lsf: one linear slip fault per solver");
  h->lsf1 = (int *)malloc((size_t)n * sizeof(int));
  h->lsf2 = (int *)malloc((size_t)n * sizeof(int));
  for (int k = 0; k < n; ++k) {
    if (node1[k] < 1 || node1[k] > h->npoin || node2[k] < 1 || node2[k] > h->npoin)
      return fail(h, 3, "set_lsf: node out of range");
    h->lsf1[k] = node1[k] - 1;
    h->lsf2[k] = node2[k] - 1;
  }
  size_t nd = (size_t)h->ndof * n;
  h->lsf_nrm = dup_d(nrm, (size_t)2 * n);
  h->lsf_B = dup_d(Bv, nd);
  h->lsf_W1 = dup_d(W1, nd);
  h->lsf_W2 = dup_d(W2, nd);
  h->lsf_kt = kt;
  h->lsf_kn = kn;
  h->nlsf = n;
  return 0;
}

int oracle_set_hht(oracle_solver *h, double alpha) {
  if (!(alpha >= 0.0 && alpha <= 1.0)) return fail(h, 3, "set_hht: alpha must lie in [0, 1]");
  if (h->it != 0) return fail(h, 4, "set_hht: the solver has already been stepped");
  if (h->leapfrog) return fail(h, 4, "set_hht: the solver is set to leapfrog");
  size_t n = (size_t)h->ndof * h->npoin;
  if (!h->da) {
    h->da = (double *)calloc(n, sizeof(double));
    h->va = (double *)calloc(n, sizeof(double));
  }
  h->hht = 1;
  h->alpha = alpha;
  return 0;
}

int oracle_set_time_scheme(oracle_solver *h, int kind) {
  if (kind != 0 && kind != 1) return fail(h, 3, "set_time_scheme: kind must be 0 (newmark) or 1 (leapfrog)");
  if (kind == 1 && (h->beta != 0.0 || h->gamma != 0.5))
    return fail(h, 3, "set_time_scheme: leapfrog needs beta = 0, gamma = 1/2 in the descriptor");
  h->leapfrog = kind;
  return 0;
}

int oracle_set_itinfo(oracle_solver *h, int itinfo) {
  if (itinfo < 0) return fail(h, 3, "set_itinfo: negative interval");
  h->itinfo = itinfo;
  return 0;
}

int oracle_field_info(oracle_solver *h, double *vmax, double *dmax, int *nonfinite) {
  double vm, dm;
  int nb;
  field_info(h, &vm, &dm, &nb);
  if (vmax) *vmax = vm;
  if (dmax) *dmax = dm;
  if (nonfinite) *nonfinite = nb;
  return 0;
}

int oracle_fetch_info(oracle_solver *h, double *rows, int *nrows) {
  if (rows && h->info_rows) memcpy(rows, h->info, sizeof(double) * 4 * h->info_rows);
  if (nrows) *nrows = h->info_rows;
  return 0;
}

int oracle_fetch_fields(oracle_solver *h, double *d, double *v, double *a) {
  size_t n = sizeof(double) * h->ndof * h->npoin;
  if (d) memcpy(d, h->d, n);
  if (v) memcpy(v, h->v, n);
  if (a) memcpy(a, h->a, n);
  return 0;
}

int oracle_set_fields(oracle_solver *h, const double *d, const double *v, const double *a) {
  size_t n = sizeof(double) * h->ndof * h->npoin;
  if (d) memcpy(h->d, d, n);
  if (v) memcpy(h->v, v, n);
  if (a) memcpy(h->a, a, n);
  return 0;
}

int oracle_fetch_seismo(oracle_solver *h, double *out, int *taken) {
  if (h->nrec) memcpy(out, h->seis, sizeof(double) * h->taken * h->ndof * h->nrec);
  if (taken) *taken = h->taken;
  return 0;
}

int oracle_fetch_fault(oracle_solver *h, double *out, int *taken) {
  if (!h->flt) return fail(h, 4, "no fault");
  memcpy(out, h->flt->rec, sizeof(double) * h->flt->taken * 6 * h->flt->nout);
  if (taken) *taken = h->flt->taken;
  return 0;
}

int oracle_fetch_fault_state(oracle_solver *h, double *D, double *V, double *T, double *MU,
                             double *Ts) {
  fault_t *f = h->flt;
  if (!f) return fail(h, 4, "no fault");
  size_t nd = (size_t)f->ndof * f->np;
  if (D) memcpy(D, f->D, nd * sizeof(double));
  if (V) memcpy(V, f->V, nd * sizeof(double));
  if (T) memcpy(T, f->T, 2 * f->np * sizeof(double));
  if (MU) memcpy(MU, f->MU, f->np * sizeof(double));
  if (Ts) memcpy(Ts, f->Tstick, 2 * f->np * sizeof(double));
  return 0;
}

int oracle_fetch_potency(oracle_solver *h, double *out, int *nrows) {
  if (!h->flt) return fail(h, 4, "no fault");
  if (out) memcpy(out, h->flt->pot, sizeof(double) * h->flt->pot_rows * 2 * (h->ndof + 1));
  if (nrows) *nrows = h->flt->pot_rows;
  return 0;
}

int oracle_get_step(oracle_solver *h, int *it) { *it = h->it; return 0; }
const char *oracle_last_error(oracle_solver *h) { return h ? h->err : "null handle"; }

void oracle_destroy(oracle_solver *h) {
  if (!h) return;
  free(h->ibool); free(h->coefs); free(h->rmass); free(h->d); free(h->v); free(h->a);
  for (int b = 0; b < h->nabso; ++b) {
    free(h->abso[b].node); free(h->abso[b].C); free(h->abso[b].enode); free(h->abso[b].K);
    free(h->abso[b].nrm);
  }
  for (int s = 0; s < h->nsrc; ++s) {
    free(h->src[s].ixi); free(h->src[s].ieta); free(h->src[s].cxi); free(h->src[s].ceta);
    free(h->src[s].ampl);
  }
  free(h->rnode); free(h->rw); free(h->seis); free(h->kv_eta); free(h->eorder); free(h->floc);
  free(h->per_m); free(h->per_s); free(h->da); free(h->va);
  free(h->lsf1); free(h->lsf2); free(h->lsf_nrm); free(h->lsf_B); free(h->lsf_W1); free(h->lsf_W2);
  for (int b = 0; b < h->nabso; ++b) {
    free(h->abso[b].wxz); free(h->abso[b].wV); free(h->abso[b].wT); free(h->abso[b].wB);
  }
  free(h->info);
  free(h->beta25);
  for (int b = 0; b < h->ndir; ++b) free(h->dir_node[b]);
  if (h->flt) {
    fault_t *f = h->flt;
    free(f->n1); free(f->n2); free(f->nrm); free(f->B); free(f->invM1); free(f->invM2);
    free(f->Z); free(f->T0); free(f->coh);
    for (int q = 0; q < 6; ++q) free(f->p[q]);
    free(f->theta); free(f->Tc); free(f->coeft); free(f->D); free(f->V); free(f->T);
    free(f->Tstick); free(f->MU); free(f->sigma); free(f->rec); free(f->srec); free(f->xz); free(f->pot); free(f);
  }
  free(h);
}
