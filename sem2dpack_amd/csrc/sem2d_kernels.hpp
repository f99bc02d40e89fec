// Device kernels of the explicit SEM time step (gfx950, wave64).
//
// Field layout: component-major SoA, f[c*npoin + node] (the reference's
// field(npoin,ndof) column-major layout, SRC/fields.f90:7-11).  R is the bulk
// type (double, or float for the FP32 bulk variant); the fault solve always
// works in double.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sem2d {

constexpr int NG = 5;  // NGLL of the structured tile kernels (OPT_NGLL, SRC/constants.f90:6)
constexpr int NP = NG * NG;
constexpr int NGMAX = 10;  // largest NGLL of the general element kernel

// h'_i(x_j) stored at H[i*5+j], and wgll(i)*wgll(j) at W2[j*5+i] (they only depend on NGLL)
__constant__ double c_H[NP];
__constant__ double c_W2[NP];

// Coefficient storage modes (sem2d_desc.coef_mode).  COEF_NODE is internal to the
// structured tile kernel: the compact moduli stored once per global node (in the field
// layout) when every element sharing a node holds the same values there.
enum { COEF_PER_ELEM = 0, COEF_SHARED = 1, COEF_COMPACT = 2, COEF_NODE = 3 };

template <typename R>
struct Coef {
  static __device__ __forceinline__ R h(int a, int b) { return (R)c_H[a * NG + b]; }
};

// ---------------------------------------------------------------------------
// Element internal forces, general (any ibool), one colour per launch.
//
// f = -K d for the elements e0 .. e0+ne-1 (consecutive in the colour-sorted
// element order), assembled into f with plain read-modify-write: elements of
// one colour share no node, so the update is race-free and deterministic.
// Follows ELAST_KD2_PSV / ELAST_KD2_SH (SRC/mat_elastic.f90:589-641, 739-775) at
// NGLL = 5 and ELAST_KD1_PSV / ELAST_KD1_SH (SRC/mat_elastic.f90:464-518, 645-681)
// otherwise:
//   dU/dxi(i,j)  = sum_k H(k,i) U(k,j)      dU/deta(i,j) = sum_k U(i,k) H(k,j)
//   f(i,j)       = sum_k H(i,k) t1(k,j) + sum_k t2(i,k) H(j,k)
// (sums in k order, as My_MATMUL and mxm form them).  One thread per GLL point,
// fe_epb(NGLL) elements per workgroup of at most 256 threads.
// ---------------------------------------------------------------------------
constexpr int fe_epb(int ngl) { return 256 / (ngl * ngl) > 0 ? 256 / (ngl * ngl) : 1; }
constexpr int FE_EPB = fe_epb(NG);
constexpr int FE_THREADS = FE_EPB * NP;

// Element coefficients a_k at point p of element e, for the three storage modes.
// Compact flat form: a(i,j,k) = W2(i,j) * modulus(i,j) * geom_e (SRC/mat_elastic.f90:327-358
// with -weights = -W2*detJ folded into geom_e).
template <typename R, int NDOF, int NELAST, int MODE, int NPL = NP>
__device__ __forceinline__ void load_coefs(const R* __restrict__ coefs, const R* __restrict__ geom, size_t e, int p,
                                           R* A) {
  static_assert(MODE != COEF_COMPACT || NPL == NP, "compact coefficients are built for NGLL = 5 only");
  if constexpr (MODE == COEF_COMPACT) {
    const R w = (R)c_W2[p];
    const R g0 = geom[(size_t)e * 3 + 0], g1 = geom[(size_t)e * 3 + 1], g2 = geom[(size_t)e * 3 + 2];
    if constexpr (NDOF == 2) {
      const R la = coefs[(size_t)e * 2 * NP + p], mu = coefs[(size_t)e * 2 * NP + NP + p];
      const R kx = la + 2 * mu;
      A[0] = w * kx * g0;
      A[1] = w * la * g2;
      A[2] = w * kx * g1;
      A[3] = w * mu * g1;
      A[4] = w * mu * g2;
      A[5] = w * mu * g0;
    } else {
      const R mu = coefs[(size_t)e * NP + p];
      A[0] = w * mu * g0;
      A[1] = w * mu * g1;
    }
  } else {
    const R* a = coefs + (MODE == COEF_SHARED ? 0 : (size_t)e * NELAST * NPL) + p;
#pragma unroll
    for (int k = 0; k < NELAST; ++k) A[k] = a[k * NPL];
  }
}

// Element forces of one colour.  With Kelvin-Voigt elements (kv != nullptr, [element
// in colour order][NGLL^2]) the element displacement is d + eta v (MAT_KV_add_etav,
// SRC/mat_kelvin_voigt.f90:139-150, applied by MAT_Fint, SRC/mat_gen.f90:435).  NGL is
// NGLL; for NGL != 5, hg holds h'_i(x_j) at hg[i*NGL+j] (for NGL = 5 it is c_H).
template <typename R, int NDOF, int NELAST, int MODE, int NGL = NG>
__global__ __launch_bounds__(256) void k_force_colour(int e0, int ne, const int* __restrict__ ibool,
                                                      const R* __restrict__ coefs, const R* __restrict__ geom,
                                                      const R* __restrict__ d, R* __restrict__ f,
                                                      int npoin, const R* __restrict__ v = nullptr,
                                                      const R* __restrict__ kv = nullptr,
                                                      const R* __restrict__ beta25 = nullptr,
                                                      const double* __restrict__ hg = nullptr) {
  constexpr int NPL = NGL * NGL;
  constexpr int EPB = fe_epb(NGL);
  // ELAST_KD1_PSV: the x force's second term is a4 dUx/deta + a5 dUz/dxi
  constexpr bool KD1 = NGL != NG;
  __shared__ R U[EPB][NDOF][NPL];
  __shared__ R T1[EPB][NDOF][NPL];
  __shared__ R T2[EPB][NDOF][NPL];
  __shared__ double sH[KD1 ? NPL : 1];
  const int t = threadIdx.x;
  if constexpr (KD1) {
    if (t < NPL) sH[t] = hg[t];
  }
  auto H = [&](int a, int b) -> R {
    if constexpr (KD1) return (R)sH[a * NGL + b];
    else return Coef<R>::h(a, b);
  };
  const int le = t / NPL;
  const int p = t - le * NPL;
  const int i = p % NGL, j = p / NGL;
  const int el = blockIdx.x * EPB + le;
  const bool active = (t < EPB * NPL) && (el < ne);
  const int e = e0 + el;
  int node = 0;
  if (active) {
    node = ibool[(size_t)e * NPL + p];
    if (kv) {
      const R eta = kv[(size_t)e * NPL + p];
#pragma unroll
      for (int c = 0; c < NDOF; ++c) U[le][c][p] = d[(size_t)c * npoin + node] + eta * v[(size_t)c * npoin + node];
    } else {
#pragma unroll
      for (int c = 0; c < NDOF; ++c) U[le][c][p] = d[(size_t)c * npoin + node];
    }
  }
  __syncthreads();
  if (active) {
    R A[NELAST];
    load_coefs<R, NDOF, NELAST, MODE, NPL>(coefs, geom, e, p, A);
    R dxi[NDOF], deta[NDOF];
#pragma unroll
    for (int c = 0; c < NDOF; ++c) {
      R s1 = 0, s2 = 0;
#pragma unroll
      for (int k = 0; k < NGL; ++k) {
        s1 += H(k, i) * U[le][c][j * NGL + k];
        s2 += U[le][c][k * NGL + i] * H(k, j);
      }
      dxi[c] = s1;
      deta[c] = s2;
    }
    if constexpr (NDOF == 2) {
      if constexpr (NELAST == 6) {
        const R a1 = A[0], a2 = A[1], a3 = A[2], a4 = A[3], a5 = A[4], a6 = A[5];
        T1[le][0][p] = a1 * dxi[0] + a2 * deta[1];
        if constexpr (KD1) T2[le][0][p] = a4 * deta[0] + a5 * dxi[1];
        else T2[le][0][p] = a4 * (deta[0] + dxi[1]);
        T1[le][1][p] = a5 * deta[0] + a6 * dxi[1];
        T2[le][1][p] = a2 * dxi[0] + a3 * deta[1];
      } else {
        T1[le][0][p] = A[0] * dxi[0] + A[6] * deta[0] + A[7] * dxi[1] + A[1] * deta[1];
        T2[le][0][p] = A[6] * dxi[0] + A[3] * deta[0] + A[4] * dxi[1] + A[8] * deta[1];
        T1[le][1][p] = A[7] * dxi[0] + A[4] * deta[0] + A[5] * dxi[1] + A[9] * deta[1];
        T2[le][1][p] = A[1] * dxi[0] + A[8] * deta[0] + A[9] * dxi[1] + A[2] * deta[1];
      }
    } else {
      if constexpr (NELAST == 2) {
        T1[le][0][p] = A[0] * dxi[0];
        T2[le][0][p] = A[1] * deta[0];
      } else {
        const R a1 = A[0], a2 = A[1], a3 = A[2];
        T1[le][0][p] = a1 * dxi[0] + a3 * deta[0];
        T2[le][0][p] = a3 * dxi[0] + a2 * deta[0];
      }
    }
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int c = 0; c < NDOF; ++c) {
      R s1 = 0, s2 = 0;
#pragma unroll
      for (int k = 0; k < NGL; ++k) {
        s1 += H(i, k) * T1[le][c][j * NGL + k];
        s2 += T2[le][c][k * NGL + i] * H(j, k);
      }
      R* fp = f + (size_t)c * npoin + node;
      if (beta25)  // 2.5D: MAT_ELAST_add_25D_f (SRC/mat_elastic.f90:447-459)
        *fp = *fp + ((s1 + s2) - beta25[(size_t)e * NPL + p] * U[le][c][p]);
      else
        *fp = *fp + (s1 + s2);
    }
  }
}

// ---------------------------------------------------------------------------
// Nodal kernels
// ---------------------------------------------------------------------------
// Newmark predictor (SRC/solver.f90:59-60) and f = 0 (f aliases a, SRC/solver.f90:52,286)
template <typename R>
__global__ void k_predict(size_t n, R* __restrict__ d, R* __restrict__ v, R* __restrict__ a, R dt,
                          R c1, R c2) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
    const R ak = a[k];
    const R vk = v[k];
    d[k] = d[k] + dt * vk + c1 * ak;
    v[k] = vk + c2 * ak;
    a[k] = 0;
  }
}

// HHT-alpha predictor (solve_HHT_alpha, SRC/solver.f90:108-113): the Newmark predictor,
// then the fields at n + alpha the forces see, da = alpha d' + (1 - alpha) d and
// va = alpha v' + (1 - alpha) v (om = 1 - alpha)
template <typename R>
__global__ void k_predict_hht(size_t n, R* __restrict__ d, R* __restrict__ v, R* __restrict__ a,
                              R* __restrict__ da, R* __restrict__ va, R dt, R c1, R c2, R al, R om) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
    const R ak = a[k], vk = v[k], dk = d[k];
    const R dn = dk + dt * vk + c1 * ak;
    const R vn = vk + c2 * ak;
    d[k] = dn;
    v[k] = vn;
    a[k] = 0;
    da[k] = al * dn + om * dk;
    va[k] = al * vn + om * vk;
  }
}

// a = f*rmass, corrector (SRC/solver.f90:78-82)
template <typename R>
__global__ void k_correct(size_t n, R* __restrict__ d, R* __restrict__ v, R* __restrict__ a,
                          const R* __restrict__ rmass, R c3, R c4) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
    const R ak = a[k] * rmass[k];
    a[k] = ak;
    v[k] = v[k] + c3 * ak;
    d[k] = d[k] + c4 * ak;
  }
}

// Point sources at step it (SRC/src_gen.f90:290-317).  Entries are applied in
// order by a single lane (a handful of nodes).
struct SrcEntry {
  int node;      // 0-based
  int comp;      // component
  int src;       // source index (row of the amplitude table)
  double coef;   // dir(c) for forces, coef_xi / coef_eta for moment stencils
};

template <typename R>
__global__ void k_sources(int nent, const SrcEntry* __restrict__ ent, const double* __restrict__ ampl,
                          int nt, int it, R* __restrict__ f, int npoin) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int q = 0; q < nent; ++q) {
    const SrcEntry s = ent[q];
    const double am = (it >= 1 && it <= nt) ? ampl[(size_t)s.src * nt + (it - 1)] : 0.0;
    R* fp = f + (size_t)s.comp * npoin + s.node;
    *fp = (R)((double)*fp + s.coef * am);
  }
}

// Absorbing boundary, one boundary per launch (SRC/bc_abso.f90:323-325)
template <typename R>
__global__ void k_abso(int n, int ndof, const int* __restrict__ node, const double* __restrict__ C,
                       const R* __restrict__ v, R* __restrict__ f, int npoin) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int g = node[k];
  for (int c = 0; c < ndof; ++c) {
    R* fp = f + (size_t)c * npoin + g;
    *fp = (R)((double)*fp - C[(size_t)c * n + k] * (double)v[(size_t)c * npoin + g]);
  }
}

// Incident plane wave through a flat absorbing boundary (BC_ABSO_apply with bc%wav,
// SRC/bc_abso.f90:300-306; SO_WAVE_get_VT, SRC/src_gen.f90:322-339): at node k the phase
// t - (x - x0) s1 - (z - z0) s2 - tdelay, the Ricker amplitude there (ricker,
// SRC/stf_ricker.f90:89-101), Vin = amp V, Tin = amp T, and the force
// f - C (v - Vin) + B Tin.  par = s1, s2, x0, z0, tdelay, f0, t0, ampli.
struct WavePar {
  double s1, s2, x0, z0, tdelay, f0, t0, ampli;
};
template <typename R>
__global__ void k_abso_wave(int n, int ndof, const int* __restrict__ node, const double* __restrict__ C,
                            const double* __restrict__ xz, const double* __restrict__ V,
                            const double* __restrict__ T, const double* __restrict__ B, WavePar W,
                            double time, const R* __restrict__ v, R* __restrict__ f, int npoin) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double phase = time - (xz[k] - W.x0) * W.s1 - (xz[n + k] - W.z0) * W.s2 - W.tdelay;
  double arg = 3.141592653589793 * W.f0 * (phase - W.t0);
  arg = arg * arg;
  const double amp = -W.ampli * (1.0 - 2.0 * arg) * exp(-arg);
  const int g = node[k];
  for (int c = 0; c < ndof; ++c) {
    R* fp = f + (size_t)c * npoin + g;
    const double vin = amp * V[(size_t)c * n + k], tin = amp * T[(size_t)c * n + k];
    *fp = (R)(((double)*fp - C[(size_t)c * n + k] * ((double)v[(size_t)c * npoin + g] - vin)) + B[k] * tin);
  }
}

// Linear slip fault (BC_LSF_apply, SRC/bc_lsf.f90:161-186): per node pair, the traction
// T = W2 f2 - W1 f1 and the jump dD = d2 - d1 (d the displacement the forces see), in the
// fault frame for P-SV (t, n from the side-1 normal), the components with finite
// stiffness replaced by k dD, back to (x, z); f1 += B T, f2 -= B T.  kt, kn = DBL_MAX:
// welded (the component is left as computed).
template <typename R>
__global__ void k_lsf(int n, int ndof, const int* __restrict__ n1, const int* __restrict__ n2,
                      const double* __restrict__ nrm, const double* __restrict__ B, const double* __restrict__ W1,
                      const double* __restrict__ W2, double kt, double kn, const R* __restrict__ d,
                      R* __restrict__ f, int npoin) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i1 = n1[k], i2 = n2[k];
  double T[2] = {0, 0}, dD[2] = {0, 0};
  for (int c = 0; c < ndof; ++c) {
    const size_t o = (size_t)c * npoin;
    T[c] = W2[(size_t)c * n + k] * (double)f[o + i2] - W1[(size_t)c * n + k] * (double)f[o + i1];
    dD[c] = (double)d[o + i2] - (double)d[o + i1];
  }
  const double big = 1.7976931348623157e308;
  if (ndof == 2) {
    const double nx = nrm[k], nz = nrm[n + k];
    double t0 = nz * T[0] - nx * T[1], t1 = nx * T[0] + nz * T[1];
    const double d0 = nz * dD[0] - nx * dD[1], d1 = nx * dD[0] + nz * dD[1];
    if (kt < big) t0 = kt * d0;
    if (kn < big) t1 = kn * d1;
    T[0] = nz * t0 + nx * t1;
    T[1] = -nx * t0 + nz * t1;
  } else if (kt < big) {
    T[0] = kt * dD[0];
  }
  for (int c = 0; c < ndof; ++c) {
    const size_t o = (size_t)c * npoin;
    const double bt = B[(size_t)c * n + k] * T[c];
    f[o + i1] = (R)((double)f[o + i1] + bt);
    f[o + i2] = (R)((double)f[o + i2] - bt);
  }
}

// Absorbing boundary that is not flat, P-SV (SRC/bc_abso.f90:313-318): with the node normal
// n, Vn = (v.n) n and the force loses C0 Vn + C1 (v - Vn), C = [rho cp, rho cs] w|J|.
template <typename R>
__global__ void k_abso_normal(int n, const int* __restrict__ node, const double* __restrict__ C,
                              const double* __restrict__ nrm, const R* __restrict__ v, R* __restrict__ f,
                              int npoin) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int g = node[k];
  const double n1 = nrm[k], n2 = nrm[n + k];
  const double v1 = (double)v[g], v2 = (double)v[npoin + g];
  const double dot = v1 * n1 + v2 * n2;
  const double vn1 = dot * n1, vn2 = dot * n2;
  f[g] = (R)((double)f[g] - C[k] * vn1 - C[n + k] * (v1 - vn1));
  f[npoin + g] = (R)((double)f[npoin + g] - C[k] * vn2 - C[n + k] * (v2 - vn2));
}

// Stacey term of one flat P-SV absorbing boundary (SRC/bc_abso.f90:321-333), general path:
// node q of the boundary (global node node[q]) subtracts KxD = sum over its boundary
// elements (t = off[q] .. off[q+1]-1, element order; idx[t] = e * ng + k) of
// K[e][k][c] * sum_j h'_j(x_k) d_c(enode[e][j]), d the predicted displacement; hg holds
// h'_i(x_j) at hg[i*ng+j].
template <typename R>
__global__ void k_stacey(int n, const int* __restrict__ node, const int* __restrict__ off,
                         const int* __restrict__ idx, const int* __restrict__ enode, const double* __restrict__ K,
                         const R* __restrict__ d, R* __restrict__ f, int npoin, int ng,
                         const double* __restrict__ hg) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int g = node[q];
  for (int c = 0; c < 2; ++c) {
    double s = 0.0;
    for (int t = off[q]; t < off[q + 1]; ++t) {
      const int ek = idx[t], e = ek / ng, k = ek - ng * e;
      double hd = 0.0;
      for (int j = 0; j < ng; ++j) hd += hg[j * ng + k] * (double)d[(size_t)c * npoin + enode[e * ng + j]];
      s += K[(size_t)ek * 2 + c] * hd;
    }
    R* fp = f + (size_t)c * npoin + g;
    *fp = (R)((double)*fp - s);
  }
}

// receivers: out[s][c][r] = field[c][node[r]]
template <typename R>
__global__ void k_receivers(int nrec, int ndof, const int* __restrict__ node, const R* __restrict__ field,
                            int npoin, double* __restrict__ out) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrec) return;
  for (int c = 0; c < ndof; ++c) out[(size_t)c * nrec + r] = (double)field[(size_t)c * npoin + node[r]];
}

// Receivers inside elements (REC_store with AtNode = F, SRC/receivers.f90:326-341): the
// np = NGLL^2 weights h_i(xi) h_j(eta) times the field at the element's nodes, summed in
// the element's (j, i) order without contraction.
#pragma clang fp contract(off)
template <typename R>
__global__ void k_receivers_interp(int nrec, int ndof, const int* __restrict__ node, const double* __restrict__ w,
                                   const R* __restrict__ field, size_t npoin, double* __restrict__ out, int np) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrec) return;
  for (int c = 0; c < ndof; ++c) {
    double s = 0.0;
    for (int k = 0; k < np; ++k) s += w[(size_t)r * np + k] * (double)field[(size_t)c * npoin + node[(size_t)r * np + k]];
    out[(size_t)c * nrec + r] = s;
  }
}
#pragma clang fp contract(on)

// BC_DIRNEU with Dirichlet components (SRC/bc_dirneu.f90:148-169): the force (MxA) of the
// boundary's nodes set to zero in the components of comp (bit c: component c)
template <typename R>
__global__ void k_dirichlet(int n, const int* __restrict__ node, int comp, int ndof, R* __restrict__ f,
                            size_t npoin) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  for (int c = 0; c < ndof; ++c)
    if (comp >> c & 1) f[(size_t)c * npoin + node[k]] = 0;
}

// Progress line of the reference's time loop (SRC/main.f90:45,73-76): max |v| and
// max |d| over the fields, and the number of non-finite values of d, v, a (the
// watchdog).  The maxima of non-negative doubles are kept as their bit patterns, whose
// integer order is their numeric order, so the result does not depend on the order in
// which waves finish.  acc: [0] max |v| bits, [1] max |d| bits; cnt: [0] non-finite
// values, [1] first step at which the check found one (0: none).
template <typename R>
__global__ void k_field_max(size_t n, const R* __restrict__ d, const R* __restrict__ v, const R* __restrict__ a,
                            unsigned long long* __restrict__ acc, int* __restrict__ cnt) {
  double vm = 0.0, dm = 0.0;
  int bad = 0;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
    const double x = (double)v[k], y = (double)d[k], z = (double)a[k];
    bad += !isfinite(x) + !isfinite(y) + !isfinite(z);
    vm = fmax(vm, fabs(x));  // fmax returns the other operand for a NaN
    dm = fmax(dm, fabs(y));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    vm = fmax(vm, __shfl_xor(vm, o));
    dm = fmax(dm, __shfl_xor(dm, o));
    bad += __shfl_xor(bad, o);
  }
  if ((threadIdx.x & 63) == 0) {
    if (vm > 0.0) atomicMax(&acc[0], (unsigned long long)__double_as_longlong(vm));
    if (dm > 0.0) atomicMax(&acc[1], (unsigned long long)__double_as_longlong(dm));
    if (bad) atomicAdd(&cnt[0], bad);
  }
}

// one progress row (it, time, max |v|, max |d|) from k_field_max's accumulators, which
// it then clears for the next row; with row set, notes the first step with a non-finite
// value (row null: a query outside the time loop, which leaves the watchdog alone)
__global__ void k_info_row(int it, double t, unsigned long long* __restrict__ acc, int* __restrict__ cnt,
                           double* __restrict__ row) {
  if (threadIdx.x != 0) return;
  if (row) {
    row[0] = (double)it;
    row[1] = t;
    row[2] = __longlong_as_double((long long)acc[0]);
    row[3] = __longlong_as_double((long long)acc[1]);
  }
  if (row && cnt[0] && !cnt[1]) cnt[1] = it > 0 ? it : -1;
  acc[0] = acc[1] = 0ull;
  cnt[0] = 0;
}

}  // namespace sem2d
