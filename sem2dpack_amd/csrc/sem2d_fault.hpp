// Dynamic fault boundary: pointwise traction solve on split-node pairs.
// asinh, sinh and cosh follow the C library's decomposition (sem2d_libm.hpp).
//
// One lane per fault node pair.  All fault arithmetic is FP64 whatever the bulk
// type, and floating-point contraction is disabled so that, given the same
// nodal inputs, every operation rounds as in the reference.
//   BC_DYNFLT_apply      SRC/bc_dynflt.f90:569-739
//   slip weakening       SRC/bc_dynflt_swf.f90:142-194
//   rate and state       SRC/bc_dynflt_rsf.f90:165-568 (two-pass rsf_solver,
//                        aging/slip/V-shape state updates, safeguarded
//                        Newton-bisection nr_solver returning the slip rate of
//                        its last function evaluation)
//   Coulomb normal law   SRC/bc_dynflt_normal.f90:118-136 (kind 1)
#pragma once
#include <hip/hip_runtime.h>

#include "sem2d_libm.hpp"

namespace sem2d {

struct FaultDev {
  int np, ndof;
  const int *n1, *n2;                    // 0-based node pairs; n2[k] = -1 on a fault with one
                                         // side (on a domain boundary, SRC/bc_dynflt.f90:256-259)
  const double *nrm;                     // [2][np]
  const double *B, *invM1, *invM2, *Z;   // [ndof][np]
  const double *T0;                      // [2][np]
  const double *coh;                     // [np]
  double a2v, a2d, dt;
  int opening, law, kind, healing;
  const double *p0, *p1, *p2, *p3, *p4, *p5;  // friction parameters [np]
  const double *Tc, *coeft;              // RSF kind 1 helpers
  double *theta;                         // [np] state
  double *D, *V;                         // [ndof][np]
  double *T, *Tstick;                    // [2][np]
  double *MU, *sigma;                    // [np]
  int *nr_fail;                          // Newton iteration overflow flag
  // structured grids: inverse mass of node 1 / node 2 of each pair [ndof][np], for the
  // corrector that k_fault applies after the solve
  const double *rmn1 = nullptr, *rmn2 = nullptr;
  // structured grids: Dirichlet components of the Dirichlet boundaries that act after
  // the fault, zeroed in the force of node 1 (bits 0-1) and node 2 (bits 2-3) before the
  // corrector (null: none)
  const unsigned char* dmask = nullptr;
  // time-weakening friction (BC_DYNFLT_TWF): kind (0: none; law 3: time weakening alone),
  // parameters MuS, MuD, Mu0, X, Z, V, L, T, Dc and the fault node coordinates [2][np]
  int twf = 0;
  double tw[9] = {};
  const double* xz = nullptr;
  // normal stress response (BC_DYNFLT_NOR): 0 cohesion (sigma stays the initial normal
  // stress), 1 Coulomb, 2 Prakash-Clifton with time scale (nor_coef = exp(-dt/T)), 3 with
  // length scale (nor_coef = dt/L, nor_v = V)
  int nor_kind = 1;
  double nor_coef = 0.0, nor_v = 0.0;
};

// normal_update (SRC/bc_dynflt_normal.f90:122-140): the normal stress the friction law
// sees, from the previous one, the trial normal traction tn and the trial slip rate dv
__host__ __device__ inline double normal_sigma(int kind, double coef, double vc, double sig_old, double tn,
                                               double dv) {
  switch (kind) {
    case 0: return sig_old;
    case 2: return tn + coef * (sig_old - tn);
    case 3: return tn + exp(-(fabs(dv) + vc) * coef) * (sig_old - tn);
    default: return tn;
  }
}

// Corrector of one split-fault node after the fault solve (a = f * rmass, then v and d;
// SRC/solver.f90:78-82), in the bulk precision and outside the no-contraction region
// below, like the corrector of every other node.
template <typename R>
__device__ __forceinline__ void fault_node_correct(R* d, R* v, R* a, size_t i, R rm, R c3, R c4) {
  const R ak = a[i] * rm;
  a[i] = ak;
  v[i] = v[i] + c3 * ak;
  d[i] = d[i] + c4 * ak;
}

#pragma clang fp contract(off)

__device__ inline double f_swf_mu(const FaultDev& F, int k) {
  const double th = F.theta[k], dc = F.p0[k], mus = F.p1[k], mud = F.p2[k];
  double mu;
  if (F.kind == 1) mu = mus - (mus - mud) * fmin(th / dc, 1.0);
  else if (F.kind == 2) mu = mud - (mud - mus) * exp(-th / dc);
  else mu = mud + (mus - mud) / pow(1.0 + th / dc, F.p4[k]);
  return mu + F.p3[k] * th;
}

// twf_mu (SRC/bc_dynflt_twf.f90:129-183) of kind 1..3 with tw = MuS, MuD, Mu0, X, Z, V, L,
// T, Dc at point (x, z), time t, slip d; kind 3 leaves the points beyond V*T unset in the
// reference: they get huge(1d0) here (no weakening).  Host and device.
__host__ __device__ inline double twf_mu(int kind, const double* tw, double x, double z, double time, double d) {
  const double mus = tw[0], mud = tw[1], mu0 = tw[2], X = tw[3], Z = tw[4];
  const double V = tw[5], L = tw[6], T = tw[7], Dc = tw[8];
  const double big = 1.7976931348623157e308;
  const double dx = x - X, dz = z - Z;
  const double dist = sqrt(dx * dx + dz * dz);
  if (kind == 1 || kind == 2) {
    double t, r;
    if (kind == 1) {
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
  }
  return big;
}

__device__ inline double f_twf_mu(const FaultDev& F, int k, double time, double d) {
  return twf_mu(F.twf, F.tw, F.xz[k], F.xz[F.np + k], time, d);
}

__device__ inline double f_rsf_mu(const FaultDev& F, int k, double v) {
  const double dc = F.p0[k], mus = F.p1[k], a = F.p2[k], b = F.p3[k], vs = F.p4[k], th = F.theta[k];
  switch (F.kind) {
    case 1: return mus + a * fabs(v) / (fabs(v) + vs) - b * th / (th + dc);
    case 2:
    case 3: return a * libm::asinh(fabs(v) / (2.0 * vs) * exp((mus + b * log(vs * th / dc)) / a));
    default: return a * libm::asinh(fabs(v) / (2.0 * vs) * exp((mus + b * log(F.p5[k] * th / dc + 1)) / a));
  }
}

__device__ inline double f_rsf_theta(const FaultDev& F, int k, double theta, double v) {
  const double dc = F.p0[k], dt = F.dt;
  switch (F.kind) {
    case 1: {
      double tn = theta * F.coeft[k] + F.Tc[k] * fabs(v) * (1.0 - F.coeft[k]);
      return tn < 1.0e-12 ? 0.0 : tn;
    }
    case 2:
    case 4: {
      const double x = fabs(v) / dc;
      const double ex = exp(-dt * x);
      if (dt * x > 1e-8) return theta * ex + (1.0 - ex) / x;
      return theta * ex + dt * (1.0 - 0.5 * dt * x);
    }
    default: {
      const double tn = dc / fabs(v);
      return tn * pow(theta / tn, exp(-dt / tn));
    }
  }
}

__device__ inline void f_fric(const FaultDev& F, int k, double theta, double tau, double tstick,
                              double sigma, double Z, double& fv, double& dfv, double& v) {
  const double dc = F.p0[k], mus = F.p1[k], a = F.p2[k], b = F.p3[k], vs = F.p4[k];
  double tmp = (F.kind == 4) ? mus + b * log(F.p5[k] * theta / dc + 1.0) : mus + b * log(vs * theta / dc);
  tmp = 2.0 * vs * exp(-tmp / a);
  v = libm::sinh(tau / (-sigma * a)) * tmp;
  fv = tstick - Z * v - tau;
  const double dv = libm::cosh(tau / (-sigma * a)) * tmp / (-sigma * a);
  dfv = -Z * dv - 1.0;
}

__device__ inline double f_nr_solver(const FaultDev& F, int k, double xL, double xR, double xacc,
                                     double theta, double tstick, double sigma, double Z, bool& ok) {
  double flow, fhigh, dfdx, v, xlow, xhigh, dx, dxold, fx;
  ok = true;
  f_fric(F, k, theta, xL, tstick, sigma, Z, flow, dfdx, v);
  f_fric(F, k, theta, xR, tstick, sigma, Z, fhigh, dfdx, v);
  double xl = xL, xr = xR;
  int guard = 0;
  while (flow * fhigh > 0 && guard++ < 2100) {
    xl = xl / 2.0;
    xr = xr * 2.0;
    f_fric(F, k, theta, xl, tstick, sigma, Z, flow, dfdx, v);
    f_fric(F, k, theta, xr, tstick, sigma, Z, fhigh, dfdx, v);
  }
  if (flow * fhigh > 0) { ok = false; return v; }
  if (flow == 0) return v;
  if (fhigh == 0) return v;
  if (flow < 0) { xlow = xl; xhigh = xr; } else { xhigh = xl; xlow = xr; }
  double xest = 0.5 * (xl + xr);
  dxold = fabs(xr - xl);
  dx = dxold;
  f_fric(F, k, theta, xest, tstick, sigma, Z, fx, dfdx, v);
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
      const double tmp = xest;
      xest = xest - dx;
      if (tmp == xest) return v;
    }
    if (fabs(dx) < fabs(xacc)) {
      f_fric(F, k, theta, xest, tstick, sigma, Z, fx, dfdx, v);
      return v;
    }
    f_fric(F, k, theta, xest, tstick, sigma, Z, fx, dfdx, v);
    if (fx < 0) xlow = xest; else xhigh = xest;
  }
  ok = false;
  return v;
}

__device__ inline double f_rsf_v(const FaultDev& F, int k, double tstick, double sigma, double theta,
                                 double Z, bool& ok) {
  ok = true;
  if (F.kind == 1) {
    const double dc = F.p0[k], mus = F.p1[k], a = F.p2[k], b = F.p3[k], vs = F.p4[k];
    // rsf_mu_no_direct reads the stored state (the state before this step), not the
    // updated estimate: SRC/bc_dynflt_rsf.f90:189-196,338
    const double th0 = F.theta[k];
    const double mnd = mus - b * th0 / (th0 + dc);
    double v = (tstick + sigma * mnd) / Z;
    const double tmp = v - vs + sigma * a / Z;
    v = 0.5 * (tmp + sqrt(tmp * tmp + 4.0 * v * vs));
    v = fmax(0.0, v);
    return v < 1.0e-12 ? 0.0 : v;
  }
  const double tol = -0.001 * F.p2[k] * sigma;
  return f_nr_solver(F, k, fmin(0.0, tstick), fmax(0.0, tstick), tol, theta, tstick, sigma, Z, ok);
}

// CORR (structured grids): the pair's two nodes are finished here too (fault_node_correct
// with the inverse masses rmn1, rmn2); the general path corrects every node afterwards.
// launched with 64 lanes per workgroup: the Newton iteration's registers are then not
// capped by the 1024-lane default (which spilled them to scratch).  time = it*dt of the
// step (SRC/main.f90:53), read by time-weakening friction only.
template <typename R, bool CORR = false>
__global__ __launch_bounds__(64) void k_fault(FaultDev F, R* __restrict__ d, R* __restrict__ v, R* __restrict__ f,
                                              int npoin, R c3 = 0, R c4 = 0, double time = 0.0) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= F.np) return;
  const int np = F.np, nd = F.ndof;
  const int i1 = F.n1[k], i2 = F.n2[k];
  double dD[2] = {0, 0}, dV[2] = {0, 0}, dA[2] = {0, 0}, T[2] = {0, 0}, Ts[2];
  for (int c = 0; c < nd; ++c) {
    const size_t o = (size_t)c * npoin;
    if (i2 >= 0) {
      dD[c] = (double)d[o + i2] - (double)d[o + i1];
      dV[c] = (double)v[o + i2] - (double)v[o + i1];
      dA[c] = F.invM2[c * np + k] * (double)f[o + i2] - F.invM1[c * np + k] * (double)f[o + i1];
    } else {  // one side: get_jump, get_weighted_jump (SRC/bc_dynflt.f90:692-719)
      dD[c] = -2.0 * (double)d[o + i1];
      dV[c] = -2.0 * (double)v[o + i1];
      dA[c] = -2.0 * F.invM1[c * np + k] * (double)f[o + i1];
    }
    T[c] = F.Z[c * np + k] * (dV[c] + F.a2v * dA[c]);
  }
  const double nx = F.nrm[k], nz = F.nrm[np + k];
  if (nd == 2) {
    double r0, r1;
    r0 = nz * dD[0] - nx * dD[1]; r1 = nx * dD[0] + nz * dD[1]; dD[0] = r0; dD[1] = r1;
    r0 = nz * dV[0] - nx * dV[1]; r1 = nx * dV[0] + nz * dV[1]; dV[0] = r0; dV[1] = r1;
    r0 = nz * dA[0] - nx * dA[1]; r1 = nx * dA[0] + nz * dA[1]; dA[0] = r0; dA[1] = r1;
    r0 = nz * T[0] - nx * T[1];   r1 = nx * T[0] + nz * T[1];   T[0] = r0;  T[1] = r1;
  }
  if (nd == 1 || i2 < 0) T[1] = 0.0;  // SRC/bc_dynflt.f90:604
  T[0] = T[0] + F.T0[k];
  T[1] = T[1] + F.T0[np + k];
  if (F.opening) T[1] = fmin(T[1], 0.0);
  const double sigma = normal_sigma(F.nor_kind, F.nor_coef, F.nor_v, F.sigma[k], T[1], dV[0]);
  F.sigma[k] = sigma;
  Ts[0] = T[0];
  Ts[1] = T[1];
  if (F.law == 2) {
    const double vold = F.V[k], Z1 = F.Z[k];
    bool ok1, ok2;
    double th = f_rsf_theta(F, k, F.theta[k], vold);
    double vn = f_rsf_v(F, k, T[0], sigma, th, Z1, ok1);
    th = f_rsf_theta(F, k, F.theta[k], 0.5 * (vold + vn));
    vn = f_rsf_v(F, k, T[0], sigma, th, Z1, ok2);
    if (!(ok1 && ok2)) atomicOr(F.nr_fail, 1);
    F.theta[k] = th;
    F.V[k] = vn;
    F.MU[k] = f_rsf_mu(F, k, vn);
    // superimposed time weakening (SRC/bc_dynflt.f90:632)
    if (F.twf) F.MU[k] = fmin(F.MU[k], f_twf_mu(F, k, time, F.D[k]));
    const double strength = -F.MU[k] * sigma;
    T[0] = copysign(fabs(strength), T[0]);
  } else if (F.law == 3) {  // time weakening alone (SRC/bc_dynflt.f90:656-657)
    F.MU[k] = f_twf_mu(F, k, time, F.D[k]);
    const double strength = F.coh[k] - F.MU[k] * sigma;
    T[0] = copysign(fmin(fabs(T[0]), strength), T[0]);
  } else {
    if (F.a2d == 0.0) {
      if (F.healing) {
        double th = F.theta[k] + fabs(dV[0]) * F.dt;
        if (fabs(dV[0]) < 1e-14) th = 0.0;
        F.theta[k] = th;
      } else {
        F.theta[k] = fabs(dD[0]);
      }
    } else {
      F.theta[k] = fabs(F.D[k]);
    }
    F.MU[k] = f_swf_mu(F, k);
    if (F.twf) F.MU[k] = fmin(F.MU[k], f_twf_mu(F, k, time, F.D[k]));  // :653
    const double strength = F.coh[k] - F.MU[k] * sigma;
    T[0] = copysign(fmin(fabs(T[0]), strength), T[0]);
  }
  T[0] = T[0] - F.T0[k];
  T[1] = T[1] - F.T0[np + k];
  Ts[0] = Ts[0] - F.T0[k];
  Ts[1] = Ts[1] - F.T0[np + k];
  F.T[k] = T[0];
  F.T[np + k] = T[1];
  F.Tstick[k] = Ts[0];
  F.Tstick[np + k] = Ts[1];
  double Tx[2] = {T[0], T[1]};
  if (nd == 2) {
    Tx[0] = nz * T[0] + nx * T[1];
    Tx[1] = -nx * T[0] + nz * T[1];
  }
  for (int c = 0; c < nd; ++c) {
    const size_t o = (size_t)c * npoin;
    const double bt = F.B[c * np + k] * Tx[c];
    f[o + i1] = (R)((double)f[o + i1] + bt);
    if (i2 >= 0) f[o + i2] = (R)((double)f[o + i2] - bt);
  }
  for (int c = 0; c < nd; ++c) {
    const double tf = T[c];
    const double da = dA[c] - tf / (F.Z[c * np + k] * F.a2v);
    F.D[c * np + k] = dD[c] + F.a2d * da;
    F.V[c * np + k] = dV[c] + F.a2v * da;
  }
  if constexpr (CORR) {
    const int dm = F.dmask ? F.dmask[k] : 0;
    for (int c = 0; c < nd; ++c) {
      const size_t o = (size_t)c * npoin;
      if (dm >> c & 1) f[o + i1] = 0;
      if (i2 >= 0 && (dm >> (2 + c) & 1)) f[o + i2] = 0;
      fault_node_correct<R>(d, v, f, o + i1, (R)F.rmn1[c * np + k], c3, c4);
      if (i2 >= 0) fault_node_correct<R>(d, v, f, o + i2, (R)F.rmn2[c * np + k], c3, c4);
    }
  }
}

// fault record: out[6][nout] from the current state (SRC/bc_dynflt.f90:762-767)
__global__ void k_fault_record(FaultDev F, int oix1, int oixd, int nout, double* __restrict__ out) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= nout) return;
  const int k = oix1 - 1 + o * oixd;
  out[0 * nout + o] = F.D[k];
  out[1 * nout + o] = F.V[k];
  out[2 * nout + o] = F.T[k];
  out[3 * nout + o] = F.T[F.np + k];
  out[4 * nout + o] = F.MU[k];
  out[5 * nout + o] = F.Tstick[k];
}

// Per-side fault record (osides = T; export_side of get_side, SRC/bc_dynflt.f90:769-812):
// out[4*ndof][nout], rows d of side 1, d of side 2, v of side 1, v of side 2, each in
// the fault frame (t, n) for P-SV (rotate, SRC/bc_dynflt.f90:722-739); side 2 of a
// fault with one side is -(side 1).  FP64, no contraction, like the fault solve.
#pragma clang fp contract(off)
template <typename R>
__global__ void k_fault_sides(FaultDev F, const R* __restrict__ d, const R* __restrict__ v, size_t stride,
                              int oix1, int oixd, int nout, double* __restrict__ out) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= nout) return;
  const int k = oix1 - 1 + o * oixd;
  const int i1 = F.n1[k], i2 = F.n2[k];
  const int nd = F.ndof;
  int row = 0;
  for (int w = 0; w < 2; ++w)
    for (int side = 1; side <= 2; ++side) {
      const R* u = w == 0 ? d : v;
      double x[2] = {0.0, 0.0};
      for (int c = 0; c < nd; ++c) {
        const R* uc = u + (size_t)c * stride;
        x[c] = side == 1 ? (double)uc[i1] : (i2 >= 0 ? (double)uc[i2] : -(double)uc[i1]);
      }
      if (nd == 1) {
        out[(size_t)row * nout + o] = x[0];
      } else {
        const double n1x = F.nrm[k], n1z = F.nrm[F.np + k];
        out[(size_t)row * nout + o] = n1z * x[0] - n1x * x[1];
        out[(size_t)(row + 1) * nout + o] = n1x * x[0] + n1z * x[1];
      }
      row += nd;
    }
}
#pragma clang fp contract(on)

// Potency of the displacement and velocity jumps (BC_DYNFLT_potency,
// SRC/bc_dynflt.f90:832-855): P-SV p11, p22, p12 of d then of v, SH p13, p23 of d
// then of v, one row of 2*(ndof+1) values, over the fault pairs k0 .. k0+n-1.  One
// workgroup of 256 lanes: lane t sums pairs k0+t, k0+t+256, ..., then a fixed-shape
// tree, so the result is run-to-run identical and the same for a fault alone (k0 = 0)
// and inside a batch (its pairs at k0 of the concatenated arrays).
template <typename R>
__device__ __forceinline__ void potency_row(const FaultDev& F, int k0, int n, const R* __restrict__ d,
                                            const R* __restrict__ v, size_t stride, double* __restrict__ out,
                                            double (*red)[256]) {
  const int t = threadIdx.x, np = F.np;
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (int kl = t; kl < n; kl += 256) {
    const int k = k0 + kl;
    const int i1 = F.n1[k], i2 = F.n2[k];
    const double nx = F.nrm[k], nz = F.nrm[np + k], b = F.B[k];
    // get_jump (SRC/bc_dynflt.f90:692-704): x2 - x1, or -2 x1 on a one-sided fault
    auto jump = [&](const R* x) {
      return i2 >= 0 ? (double)x[i2] - (double)x[i1] : -2.0 * (double)x[i1];
    };
    const double jd0 = jump(d), jv0 = jump(v);
    if (F.ndof == 2) {
      const double jd1 = jump(d + stride);
      const double jv1 = jump(v + stride);
      s[0] += nx * jd0 * b;
      s[1] += nz * jd1 * b;
      s[2] += (nx * jd1 + nz * jd0) * b;
      s[3] += nx * jv0 * b;
      s[4] += nz * jv1 * b;
      s[5] += (nx * jv1 + nz * jv0) * b;
    } else {
      s[0] += nx * jd0 * b;
      s[1] += nz * jd0 * b;
      s[2] += nx * jv0 * b;
      s[3] += nz * jv0 * b;
    }
  }
  for (int q = 0; q < 6; ++q) red[q][t] = s[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w)
      for (int q = 0; q < 6; ++q) red[q][t] = red[q][t] + red[q][t + w];
    __syncthreads();
  }
  if (t == 0) {
    if (F.ndof == 2) {
      out[0] = red[0][0];
      out[1] = red[1][0];
      out[2] = 0.5 * red[2][0];
      out[3] = red[3][0];
      out[4] = red[4][0];
      out[5] = 0.5 * red[5][0];
    } else {
      out[0] = 0.5 * red[0][0];
      out[1] = 0.5 * red[1][0];
      out[2] = 0.5 * red[2][0];
      out[3] = 0.5 * red[3][0];
    }
  }
}

template <typename R>
__global__ __launch_bounds__(256) void k_fault_potency(FaultDev F, const R* __restrict__ d, const R* __restrict__ v,
                                                       size_t stride, double* __restrict__ out) {
  __shared__ double red[6][256];
  potency_row<R>(F, 0, F.np, d, v, stride, out, red);
}

// Batched potency: workgroup m writes member m's row for the state after the current
// step, row row_shift + *itp of pot[m][cap][2*(ndof+1)] (itp: the batch's step counter
// in device memory, which a replayed graph advances); fbase[m] .. fbase[m+1] are member
// m's pairs in the concatenated fault arrays.
template <typename R>
__global__ __launch_bounds__(256) void k_batch_potency(FaultDev F, const int* __restrict__ fbase,
                                                       const R* __restrict__ d, const R* __restrict__ v,
                                                       size_t stride, const int* __restrict__ itp, int row_shift,
                                                       int cap, double* __restrict__ pot) {
  __shared__ double red[6][256];
  const int m = blockIdx.x;
  const int row = row_shift + *itp;
  if (row < 0 || row >= cap) return;  // uniform over the workgroup; the host reserves the rows
  const int k0 = fbase[m];
  potency_row<R>(F, k0, fbase[m + 1] - k0, d, v, stride, pot + ((size_t)m * cap + row) * (2 * (F.ndof + 1)), red);
}

#pragma clang fp contract(on)

}  // namespace sem2d
