// sem2d_solver.hip -- C ABI (include/sem2d.h) of the MI355X explicit SEM time step.
//
// Host side: owns the device buffers of one problem and one HIP stream, builds
// the element colour schedule, and launches the per-step kernel sequence of
// solve_Newmark (SRC/solver.f90:42-84):
//   predictor -> element forces (one launch per colour) -> sources -> absorbing
//   boundaries (in BC order) -> fault -> a = f*rmass + corrector -> receiver and
//   fault records.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "sem2d.h"
#include "sem2d_fault.hpp"
#include "sem2d_kernels.hpp"

using namespace sem2d;

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct AbsoDev {
  int n;
  int* node;
  double* C;
};

struct FaultHost {
  bool on = false;
  FaultDev dev;
  std::vector<void*> allocs;
  int oit, oitd, oix1, oixn, oixd, nout, nsamp, taken;
  double* rec;  // device [nsamp][6][nout]
};

}  // namespace

struct sem2d_solver {
  int ndof, nelem, npoin, nelast, coef_mode, fp32;
  double dt, beta, gamma;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // bulk arrays (type R = double or float, stored as raw bytes)
  void *d = nullptr, *v = nullptr, *a = nullptr, *rmass = nullptr, *coefs = nullptr, *geom = nullptr;
  int* ibool = nullptr;  // colour-sorted, 0-based
  std::vector<int> colour_start;  // ncolour+1 offsets into the sorted order
  std::vector<int> elem_order;    // sorted position -> original element (0-based)
  // boundaries and sources
  std::vector<AbsoDev> abso;
  std::vector<SrcEntry> src_host;
  std::vector<std::vector<double>> stf;
  int nsrc = 0;
  SrcEntry* src_dev = nullptr;
  double* ampl_dev = nullptr;
  int ampl_nt = 0;
  bool src_dirty = false;
  // receivers
  int nrec = 0, isamp = 1, field = 1, nsamp = 0, taken = 0;
  int* rec_node = nullptr;
  double* seis = nullptr;
  FaultHost flt;
  int* flags = nullptr;  // [0] Newton failure
  int it = 0;
  double last_ms = 0.0;
  std::string err;
  std::vector<void*> allocs;
};

namespace {

int set_err(sem2d_solver* h, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (h) h->err = buf;
  return code;
}

#define HIPCHK(h, call)                                                                   \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return set_err(h, 10, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

template <typename T>
int dev_alloc(sem2d_solver* h, T** p, size_t n) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) return set_err(h, 11, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
  h->allocs.push_back(q);
  *p = (T*)q;
  return 0;
}

template <typename T>
int upload(sem2d_solver* h, T** p, const T* src, size_t n) {
  int rc = dev_alloc(h, p, n);
  if (rc) return rc;
  if (n) HIPCHK(h, hipMemcpyAsync(*p, src, n * sizeof(T), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

size_t rsize(const sem2d_solver* h) { return h->fp32 ? sizeof(float) : sizeof(double); }

// copy a double host array into a bulk (R) device array
int upload_bulk(sem2d_solver* h, void** p, const double* src, size_t n) {
  if (!h->fp32) return upload(h, (double**)p, src, n);
  std::vector<float> tmp(n);
  for (size_t k = 0; k < n; ++k) tmp[k] = (float)src[k];
  return upload(h, (float**)p, tmp.data(), n);
}

int write_bulk(sem2d_solver* h, void* dst, const double* src, size_t n) {
  if (!h->fp32) {
    HIPCHK(h, hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
  }
  std::vector<float> tmp(n);
  for (size_t k = 0; k < n; ++k) tmp[k] = (float)src[k];
  HIPCHK(h, hipMemcpyAsync(dst, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int read_bulk(sem2d_solver* h, double* dst, const void* src, size_t n) {
  if (!h->fp32) {
    HIPCHK(h, hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
  }
  std::vector<float> tmp(n);
  HIPCHK(h, hipMemcpyAsync(tmp.data(), src, n * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (size_t k = 0; k < n; ++k) dst[k] = tmp[k];
  return 0;
}

// Element colouring: no two elements of one colour share a node.
// Structured boxes use the 2x2 checkerboard of cells; otherwise greedy first-fit
// in element order.
int build_colours(sem2d_solver* h, const sem2d_desc* s, std::vector<int>& colour, int& ncol) {
  const int ne = s->nelem;
  colour.assign(ne, -1);
  if (s->nx > 0 && s->elem_of_cell) {
    if ((long)s->nx * s->nz != ne) return set_err(h, 3, "nx*nz (%d*%d) != nelem (%d)", s->nx, s->nz, ne);
    for (int iz = 0; iz < s->nz; ++iz)
      for (int ix = 0; ix < s->nx; ++ix) {
        int e = s->elem_of_cell[(size_t)iz * s->nx + ix] - 1;
        if (e < 0 || e >= ne || colour[e] != -1) return set_err(h, 3, "elem_of_cell is not a permutation");
        colour[e] = (ix & 1) + 2 * (iz & 1);
      }
    ncol = 4;
  } else {
    std::vector<std::vector<int>> node_elems(s->npoin);
    for (int e = 0; e < ne; ++e)
      for (int p = 0; p < NP; ++p) node_elems[s->ibool[(size_t)e * NP + p] - 1].push_back(e);
    ncol = 0;
    std::vector<int> used;
    for (int e = 0; e < ne; ++e) {
      used.assign(ncol + 1, 0);
      for (int p = 0; p < NP; ++p)
        for (int o : node_elems[s->ibool[(size_t)e * NP + p] - 1])
          if (colour[o] >= 0) used[colour[o]] = 1;
      int c = 0;
      while (c < ncol && used[c]) ++c;
      colour[e] = c;
      if (c == ncol) ++ncol;
    }
  }
  // verify: each node is touched at most once per colour
  std::vector<int> stamp((size_t)s->npoin, -1);
  for (int c = 0; c < ncol; ++c)
    for (int e = 0; e < ne; ++e) {
      if (colour[e] != c) continue;
      for (int p = 0; p < NP; ++p) {
        int g = s->ibool[(size_t)e * NP + p] - 1;
        if (stamp[g] == c) {
          // repeated node inside one element is legal only for the element itself
          bool same = false;
          for (int q = 0; q < p; ++q)
            if (s->ibool[(size_t)e * NP + q] - 1 == g) same = true;
          if (!same) return set_err(h, 3, "colouring conflict at node %d colour %d", g + 1, c);
        }
        stamp[g] = c;
      }
    }
  return 0;
}

template <typename R, int NDOF, int NELAST, int MODE>
void launch_force_t(sem2d_solver* h, const R* d, R* f) {
  for (size_t c = 0; c + 1 < h->colour_start.size(); ++c) {
    const int e0 = h->colour_start[c], ne = h->colour_start[c + 1] - e0;
    if (ne <= 0) continue;
    k_force_colour<R, NDOF, NELAST, MODE><<<(ne + FE_EPB - 1) / FE_EPB, 256, 0, h->stream>>>(
        e0, ne, h->ibool, (const R*)h->coefs, (const R*)h->geom, d, f, h->npoin);
  }
}

template <typename R, int NDOF, int NELAST>
void launch_force_m(sem2d_solver* h, const R* d, R* f) {
  switch (h->coef_mode) {
    case COEF_SHARED: launch_force_t<R, NDOF, NELAST, COEF_SHARED>(h, d, f); break;
    case COEF_COMPACT:
      if constexpr (NELAST == 2 || NELAST == 6) launch_force_t<R, NDOF, NELAST, COEF_COMPACT>(h, d, f);
      break;
    default: launch_force_t<R, NDOF, NELAST, COEF_PER_ELEM>(h, d, f);
  }
}

template <typename R>
void launch_force(sem2d_solver* h, const R* d, R* f) {
  if (h->ndof == 2) {
    if (h->nelast == 6) launch_force_m<R, 2, 6>(h, d, f);
    else launch_force_m<R, 2, 10>(h, d, f);
  } else {
    if (h->nelast == 2) launch_force_m<R, 1, 2>(h, d, f);
    else launch_force_m<R, 1, 3>(h, d, f);
  }
}

int grid_for(size_t n, int block) {
  size_t g = (n + block - 1) / block;
  return (int)std::min<size_t>(std::max<size_t>(g, 1), 65536);
}

int sync_sources(sem2d_solver* h) {
  if (!h->src_dirty) return 0;
  h->src_dirty = false;
  int nt = 0;
  for (auto& t : h->stf) nt = std::max<int>(nt, (int)t.size());
  std::vector<double> tab((size_t)std::max(h->nsrc, 1) * std::max(nt, 1), 0.0);
  for (int s = 0; s < h->nsrc; ++s)
    for (size_t k = 0; k < h->stf[s].size(); ++k) tab[(size_t)s * nt + k] = h->stf[s][k];
  if (h->src_dev) hipFree(h->src_dev);
  if (h->ampl_dev) hipFree(h->ampl_dev);
  h->src_dev = nullptr;
  h->ampl_dev = nullptr;
  HIPCHK(h, hipMalloc(&h->src_dev, std::max<size_t>(h->src_host.size(), 1) * sizeof(SrcEntry)));
  HIPCHK(h, hipMalloc(&h->ampl_dev, tab.size() * sizeof(double)));
  if (!h->src_host.empty())
    HIPCHK(h, hipMemcpyAsync(h->src_dev, h->src_host.data(), h->src_host.size() * sizeof(SrcEntry),
                             hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->ampl_dev, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->ampl_nt = nt;
  return 0;
}

template <typename R>
int record_receivers(sem2d_solver* h) {
  if (!h->nrec || h->it % h->isamp) return 0;
  int s = h->it / h->isamp;
  if (s >= h->nsamp) return 0;
  const R* fld = (const R*)(h->field == 0 ? h->d : (h->field == 1 ? h->v : h->a));
  k_receivers<R><<<grid_for(h->nrec, 64), 64, 0, h->stream>>>(h->nrec, h->ndof, h->rec_node, fld, h->npoin,
                                                              h->seis + (size_t)s * h->ndof * h->nrec);
  h->taken = std::max(h->taken, s + 1);
  return 0;
}

int record_fault(sem2d_solver* h) {
  FaultHost& F = h->flt;
  if (!F.on || h->it < F.oit) return 0;
  if (F.taken < F.nsamp) {
    k_fault_record<<<grid_for(F.nout, 64), 64, 0, h->stream>>>(F.dev, F.oix1, F.oixd, F.nout,
                                                               F.rec + (size_t)F.taken * 6 * F.nout);
    F.taken++;
  }
  F.oit += F.oitd;
  return 0;
}

template <typename R>
int do_steps(sem2d_solver* h, int nsteps) {
  int rc = sync_sources(h);
  if (rc) return rc;
  const size_t n = (size_t)h->ndof * h->npoin;
  R* d = (R*)h->d;
  R* v = (R*)h->v;
  R* a = (R*)h->a;
  const R dt = (R)h->dt;
  const R c1 = (R)((0.5 - h->beta) * h->dt * h->dt), c2 = (R)((1.0 - h->gamma) * h->dt);
  const R c3 = (R)(h->gamma * h->dt), c4 = (R)(h->beta * h->dt * h->dt);
  const int gb = grid_for(n, 256);
  HIPCHK(h, hipEventRecord(h->ev0, h->stream));
  for (int s = 0; s < nsteps; ++s) {
    const int it = h->it + 1;
    k_predict<R><<<gb, 256, 0, h->stream>>>(n, d, v, a, dt, c1, c2);
    launch_force<R>(h, d, a);
    if (!h->src_host.empty())
      k_sources<R><<<1, 64, 0, h->stream>>>((int)h->src_host.size(), h->src_dev, h->ampl_dev, h->ampl_nt, it, a,
                                           h->npoin);
    for (auto& b : h->abso)
      k_abso<R><<<grid_for(b.n, 128), 128, 0, h->stream>>>(b.n, h->ndof, b.node, b.C, v, a, h->npoin);
    if (h->flt.on) k_fault<R><<<grid_for(h->flt.dev.np, 64), 64, 0, h->stream>>>(h->flt.dev, d, v, a, h->npoin);
    k_correct<R><<<gb, 256, 0, h->stream>>>(n, d, v, a, (const R*)h->rmass, c3, c4);
    h->it = it;
    record_receivers<R>(h);
    record_fault(h);
  }
  HIPCHK(h, hipEventRecord(h->ev1, h->stream));
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last_ms = nsteps > 0 ? ms / nsteps : 0.0;
  if (h->flt.on) {
    int flag = 0;
    HIPCHK(h, hipMemcpy(&flag, h->flags, sizeof(int), hipMemcpyDeviceToHost));
    if (flag) return set_err(h, 20, "NR_Solver has exceeded the maximum iterations (200)");
  }
  return 0;
}

}  // namespace

extern "C" {

int sem2d_create(const sem2d_desc* s, sem2d_solver** out) {
  if (!s || !out) return 1;
  *out = nullptr;
  if (s->ngll != NG) return 2;
  if (!(s->ndof == 1 || s->ndof == 2)) return 2;
  if (s->ndof == 2 && s->nelast != 6 && s->nelast != 10) return 2;
  if (s->ndof == 1 && s->nelast != 2 && s->nelast != 3) return 2;
  if (s->nelem <= 0 || s->npoin <= 0 || !s->ibool || !s->hprime || !s->rmass) return 2;
  if (s->coef_mode < 0 || s->coef_mode > 2) return 2;
  if (s->coef_mode == COEF_COMPACT && (!s->lam_mu || !s->geom || !s->wgll || (s->nelast != 2 && s->nelast != 6)))
    return 2;
  if (s->coef_mode != COEF_COMPACT && !s->coefs) return 2;
  sem2d_solver* h = new sem2d_solver();
  *out = h;
  h->ndof = s->ndof;
  h->nelem = s->nelem;
  h->npoin = s->npoin;
  h->nelast = s->nelast;
  h->coef_mode = s->coef_mode;
  h->fp32 = s->bulk_fp32 ? 1 : 0;
  h->dt = s->dt;
  h->beta = s->beta;
  h->gamma = s->gamma;
  for (size_t k = 0; k < (size_t)s->nelem * NP; ++k)
    if (s->ibool[k] < 1 || s->ibool[k] > s->npoin) return set_err(h, 3, "ibool entry %zu out of range", k);
  HIPCHK(h, hipStreamCreate(&h->stream));
  HIPCHK(h, hipEventCreate(&h->ev0));
  HIPCHK(h, hipEventCreate(&h->ev1));
  double Hc[NP];
  for (int i = 0; i < NG; ++i)
    for (int j = 0; j < NG; ++j) Hc[i * NG + j] = s->hprime[j * NG + i];
  HIPCHK(h, hipMemcpyToSymbol(HIP_SYMBOL(c_H), Hc, sizeof Hc, 0, hipMemcpyHostToDevice));
  if (s->coef_mode == COEF_COMPACT) {
    double W2[NP];
    for (int j = 0; j < NG; ++j)
      for (int i = 0; i < NG; ++i) W2[j * NG + i] = s->wgll[i] * s->wgll[j];
    HIPCHK(h, hipMemcpyToSymbol(HIP_SYMBOL(c_W2), W2, sizeof W2, 0, hipMemcpyHostToDevice));
  }

  std::vector<int> colour;
  int ncol = 0;
  int rc = build_colours(h, s, colour, ncol);
  if (rc) return rc;
  h->colour_start.assign(ncol + 1, 0);
  for (int e = 0; e < s->nelem; ++e) h->colour_start[colour[e] + 1]++;
  for (int c = 0; c < ncol; ++c) h->colour_start[c + 1] += h->colour_start[c];
  std::vector<int> pos(h->colour_start.begin(), h->colour_start.end() - 1);
  h->elem_order.assign(s->nelem, 0);
  for (int e = 0; e < s->nelem; ++e) h->elem_order[pos[colour[e]]++] = e;
  std::vector<int> ib((size_t)s->nelem * NP);
  for (int k = 0; k < s->nelem; ++k) {
    int e = h->elem_order[k];
    for (int p = 0; p < NP; ++p) ib[(size_t)k * NP + p] = s->ibool[(size_t)e * NP + p] - 1;
  }
  if ((rc = upload(h, &h->ibool, ib.data(), ib.size()))) return rc;
  if (h->coef_mode == COEF_SHARED) {
    if ((rc = upload_bulk(h, &h->coefs, s->coefs, (size_t)h->nelast * NP))) return rc;
  } else {
    // per-element arrays follow the colour-sorted element order
    const size_t per = h->coef_mode == COEF_COMPACT ? (size_t)h->ndof * NP : (size_t)h->nelast * NP;
    const double* src = h->coef_mode == COEF_COMPACT ? s->lam_mu : s->coefs;
    std::vector<double> co((size_t)s->nelem * per);
    for (int k = 0; k < s->nelem; ++k)
      std::memcpy(&co[(size_t)k * per], src + (size_t)h->elem_order[k] * per, per * sizeof(double));
    if ((rc = upload_bulk(h, &h->coefs, co.data(), co.size()))) return rc;
    if (h->coef_mode == COEF_COMPACT) {
      std::vector<double> ge((size_t)s->nelem * 3);
      for (int k = 0; k < s->nelem; ++k)
        for (int q = 0; q < 3; ++q) ge[(size_t)k * 3 + q] = s->geom[(size_t)h->elem_order[k] * 3 + q];
      if ((rc = upload_bulk(h, &h->geom, ge.data(), ge.size()))) return rc;
    }
  }
  const size_t nf = (size_t)h->ndof * h->npoin;
  if ((rc = upload_bulk(h, &h->rmass, s->rmass, nf))) return rc;
  if ((rc = dev_alloc(h, (char**)&h->d, nf * rsize(h)))) return rc;
  if ((rc = dev_alloc(h, (char**)&h->v, nf * rsize(h)))) return rc;
  if ((rc = dev_alloc(h, (char**)&h->a, nf * rsize(h)))) return rc;
  if ((rc = dev_alloc(h, &h->flags, 4))) return rc;
  HIPCHK(h, hipMemsetAsync(h->d, 0, nf * rsize(h), h->stream));
  HIPCHK(h, hipMemsetAsync(h->v, 0, nf * rsize(h), h->stream));
  HIPCHK(h, hipMemsetAsync(h->a, 0, nf * rsize(h), h->stream));
  HIPCHK(h, hipMemsetAsync(h->flags, 0, 4 * sizeof(int), h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int sem2d_set_abso(sem2d_solver* h, int n, const int* node, const double* C) {
  if (!h) return 1;
  if (n <= 0 || !node || !C) return set_err(h, 3, "set_abso: empty boundary");
  std::vector<int> nd(n);
  for (int k = 0; k < n; ++k) {
    if (node[k] < 1 || node[k] > h->npoin) return set_err(h, 3, "set_abso: node out of range");
    nd[k] = node[k] - 1;
  }
  AbsoDev b;
  b.n = n;
  int rc = upload(h, &b.node, nd.data(), nd.size());
  if (rc) return rc;
  if ((rc = upload(h, &b.C, C, (size_t)h->ndof * n))) return rc;
  h->abso.push_back(b);
  return 0;
}

int sem2d_set_force(sem2d_solver* h, int iglob, const double* dir, int* isrc) {
  if (!h) return 1;
  if (iglob < 1 || iglob > h->npoin) return set_err(h, 3, "set_force: node out of range");
  const int s = h->nsrc++;
  if (h->ndof == 1) {
    h->src_host.push_back({iglob - 1, 0, s, 1.0});
  } else {
    if (!dir) return set_err(h, 3, "set_force: dir required for P-SV");
    h->src_host.push_back({iglob - 1, 0, s, dir[0]});
    h->src_host.push_back({iglob - 1, 1, s, dir[1]});
  }
  h->stf.emplace_back();
  h->src_dirty = true;
  if (isrc) *isrc = s;
  return 0;
}

int sem2d_set_moment(sem2d_solver* h, int nel, const int* ixi, const int* ieta, const double* cxi,
                     const double* ceta, int* isrc) {
  if (!h) return 1;
  if (nel <= 0 || !ixi || !ieta || !cxi || !ceta) return set_err(h, 3, "set_moment: empty stencil");
  const int s = h->nsrc++;
  for (int k = 0; k < nel; ++k) {
    // order of SRC/src_moment.f90:229-232: xi line (all components), then eta line
    for (int c = 0; c < h->ndof; ++c)
      for (int q = 0; q < NG; ++q) {
        int g = ixi[k * NG + q];
        if (g < 1 || g > h->npoin) return set_err(h, 3, "set_moment: node out of range");
        h->src_host.push_back({g - 1, c, s, cxi[((size_t)k * h->ndof + c) * NG + q]});
      }
    for (int c = 0; c < h->ndof; ++c)
      for (int q = 0; q < NG; ++q) {
        int g = ieta[k * NG + q];
        if (g < 1 || g > h->npoin) return set_err(h, 3, "set_moment: node out of range");
        h->src_host.push_back({g - 1, c, s, ceta[((size_t)k * h->ndof + c) * NG + q]});
      }
  }
  h->stf.emplace_back();
  h->src_dirty = true;
  if (isrc) *isrc = s;
  return 0;
}

int sem2d_set_stf(sem2d_solver* h, int isrc, int nt, const double* ampl) {
  if (!h) return 1;
  if (isrc < 0 || isrc >= h->nsrc) return set_err(h, 3, "set_stf: bad source index %d", isrc);
  if (nt < 0 || (nt > 0 && !ampl)) return set_err(h, 3, "set_stf: bad table");
  h->stf[isrc].assign(ampl, ampl + nt);
  h->src_dirty = true;
  return 0;
}

int sem2d_set_receivers(sem2d_solver* h, int nrec, const int* iglob, int isamp, int field, int nsamp) {
  if (!h) return 1;
  if (nrec <= 0 || !iglob || isamp < 1 || field < 0 || field > 2 || nsamp < 1)
    return set_err(h, 3, "set_receivers: bad arguments");
  std::vector<int> nd(nrec);
  for (int r = 0; r < nrec; ++r) {
    if (iglob[r] < 1 || iglob[r] > h->npoin) return set_err(h, 3, "set_receivers: node out of range");
    nd[r] = iglob[r] - 1;
  }
  int rc = upload(h, &h->rec_node, nd.data(), nd.size());
  if (rc) return rc;
  if ((rc = dev_alloc(h, &h->seis, (size_t)nsamp * h->ndof * nrec))) return rc;
  h->nrec = nrec;
  h->isamp = isamp;
  h->field = field;
  h->nsamp = nsamp;
  h->taken = 0;
  if (h->fp32) rc = record_receivers<float>(h);
  else rc = record_receivers<double>(h);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return rc;
}

int sem2d_set_fault(sem2d_solver* h, const sem2d_fault_desc* f, int nsamp) {
  if (!h) return 1;
  if (!f || f->npoin <= 0 || nsamp < 1) return set_err(h, 3, "set_fault: bad arguments");
  if (h->flt.on) return set_err(h, 3, "set_fault: only one fault per solver");
  if (f->law != 1 && f->law != 2) return set_err(h, 3, "set_fault: law must be 1 (SWF) or 2 (RSF)");
  if ((f->law == 1 && (f->kind < 1 || f->kind > 3)) || (f->law == 2 && (f->kind < 1 || f->kind > 4)))
    return set_err(h, 3, "set_fault: invalid friction kind");
  if (f->oitd < 1 || f->oixd < 1 || f->oix1 < 1 || f->oixn > f->npoin || f->oixn < f->oix1)
    return set_err(h, 3, "set_fault: bad output decimation");
  const int np = f->npoin, nd = h->ndof;
  FaultHost& F = h->flt;
  FaultDev& D = F.dev;
  D.np = np;
  D.ndof = nd;
  std::vector<int> n1(np), n2(np);
  for (int k = 0; k < np; ++k) {
    if (f->node1[k] < 1 || f->node1[k] > h->npoin || f->node2[k] < 1 || f->node2[k] > h->npoin)
      return set_err(h, 3, "set_fault: node out of range");
    n1[k] = f->node1[k] - 1;
    n2[k] = f->node2[k] - 1;
  }
  int rc;
  int* ip;
  double* dp;
  if ((rc = upload(h, &ip, n1.data(), np))) return rc;
  D.n1 = ip;
  if ((rc = upload(h, &ip, n2.data(), np))) return rc;
  D.n2 = ip;
#define UP(field, src, n)                           \
  if ((rc = upload(h, &dp, (src), (size_t)(n)))) return rc; \
  D.field = dp;
  UP(nrm, f->n1, 2 * np);
  UP(B, f->B, nd * np);
  UP(invM1, f->invM1, nd * np);
  UP(invM2, f->invM2, nd * np);
  UP(Z, f->Z, nd * np);
  UP(T0, f->T0, 2 * np);
  UP(coh, f->cohesion, np);
  std::vector<double> zero((size_t)2 * np, 0.0);
  const double* pp[6];
  for (int q = 0; q < 6; ++q) pp[q] = f->p[q] ? f->p[q] : zero.data();
  UP(p0, pp[0], np);
  UP(p1, pp[1], np);
  UP(p2, pp[2], np);
  UP(p3, pp[3], np);
  UP(p4, pp[4], np);
  UP(p5, pp[5], np);
  std::vector<double> Tc(np, 0.0), coeft(np, 0.0), mu(np), sig(np), th(np);
  for (int k = 0; k < np; ++k) th[k] = f->theta0 ? f->theta0[k] : 0.0;
  if (f->law == 2)
    for (int k = 0; k < np; ++k) {
      Tc[k] = pp[0][k] / pp[4][k];
      coeft[k] = std::exp(-h->dt / Tc[k]);
    }
  UP(Tc, Tc.data(), np);
  UP(coeft, coeft.data(), np);
  double* theta;
  if ((rc = upload(h, &theta, th.data(), np))) return rc;
  D.theta = theta;
  std::vector<double> V0((size_t)nd * np, 0.0);
  if (f->V0)
    for (int k = 0; k < np; ++k) V0[k] = f->V0[k];
  double *Dd, *Vd, *Td, *Tsd, *MUd, *Sd;
  if ((rc = upload(h, &Dd, zero.data(), (size_t)nd * np))) return rc;
  if ((rc = upload(h, &Vd, V0.data(), (size_t)nd * np))) return rc;
  if ((rc = upload(h, &Td, zero.data(), (size_t)2 * np))) return rc;
  if ((rc = upload(h, &Tsd, zero.data(), (size_t)2 * np))) return rc;
  // initial friction coefficient (SRC/bc_dynflt.f90:404-413), computed on the host
  for (int k = 0; k < np; ++k) {
    sig[k] = f->T0[np + k];
    const double dc = pp[0][k], mus = pp[1][k];
    if (f->law == 1) {
      const double mud = pp[2][k], t = th[k];
      double m;
      if (f->kind == 1) m = mus - (mus - mud) * std::fmin(t / dc, 1.0);
      else if (f->kind == 2) m = mud - (mud - mus) * std::exp(-t / dc);
      else m = mud + (mus - mud) / std::pow(1.0 + t / dc, pp[4][k]);
      mu[k] = m + pp[3][k] * t;
    } else {
      const double a = pp[2][k], b = pp[3][k], vs = pp[4][k], v = V0[k], t = th[k];
      if (f->kind == 1) mu[k] = mus + a * std::fabs(v) / (std::fabs(v) + vs) - b * t / (t + dc);
      else if (f->kind <= 3)
        mu[k] = a * std::asinh(std::fabs(v) / (2.0 * vs) * std::exp((mus + b * std::log(vs * t / dc)) / a));
      else
        mu[k] = a * std::asinh(std::fabs(v) / (2.0 * vs) *
                               std::exp((mus + b * std::log(pp[5][k] * t / dc + 1)) / a));
    }
  }
  if ((rc = upload(h, &MUd, mu.data(), np))) return rc;
  if ((rc = upload(h, &Sd, sig.data(), np))) return rc;
#undef UP
  D.D = Dd;
  D.V = Vd;
  D.T = Td;
  D.Tstick = Tsd;
  D.MU = MUd;
  D.sigma = Sd;
  D.a2v = f->coefA2V;
  D.a2d = f->coefA2D;
  D.dt = h->dt;
  D.opening = f->allow_opening;
  D.law = f->law;
  D.kind = f->kind;
  D.healing = f->healing;
  D.nr_fail = h->flags;
  F.oit = f->oit;
  F.oitd = f->oitd;
  F.oix1 = f->oix1;
  F.oixn = f->oixn;
  F.oixd = f->oixd;
  F.nout = (f->oixn - f->oix1) / f->oixd + 1;
  F.nsamp = nsamp;
  F.taken = 0;
  if ((rc = dev_alloc(h, &F.rec, (size_t)nsamp * 6 * F.nout))) return rc;
  F.on = true;
  record_fault(h);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int sem2d_step(sem2d_solver* h, int nsteps) {
  if (!h) return 1;
  if (nsteps < 0) return set_err(h, 3, "step: negative nsteps");
  return h->fp32 ? do_steps<float>(h, nsteps) : do_steps<double>(h, nsteps);
}

int sem2d_fetch_fields(sem2d_solver* h, double* d, double* v, double* a) {
  if (!h) return 1;
  const size_t n = (size_t)h->ndof * h->npoin;
  int rc = 0;
  if (d && (rc = read_bulk(h, d, h->d, n))) return rc;
  if (v && (rc = read_bulk(h, v, h->v, n))) return rc;
  if (a && (rc = read_bulk(h, a, h->a, n))) return rc;
  return 0;
}

int sem2d_set_fields(sem2d_solver* h, const double* d, const double* v, const double* a) {
  if (!h) return 1;
  const size_t n = (size_t)h->ndof * h->npoin;
  int rc = 0;
  if (d && (rc = write_bulk(h, h->d, d, n))) return rc;
  if (v && (rc = write_bulk(h, h->v, v, n))) return rc;
  if (a && (rc = write_bulk(h, h->a, a, n))) return rc;
  return 0;
}

int sem2d_fetch_seismo(sem2d_solver* h, double* out, int* taken) {
  if (!h) return 1;
  if (h->nrec && out)
    HIPCHK(h, hipMemcpy(out, h->seis, (size_t)h->taken * h->ndof * h->nrec * sizeof(double), hipMemcpyDeviceToHost));
  if (taken) *taken = h->taken;
  return 0;
}

int sem2d_fetch_fault(sem2d_solver* h, double* out, int* taken) {
  if (!h) return 1;
  if (!h->flt.on) return set_err(h, 4, "fetch_fault: no fault");
  if (out)
    HIPCHK(h, hipMemcpy(out, h->flt.rec, (size_t)h->flt.taken * 6 * h->flt.nout * sizeof(double),
                        hipMemcpyDeviceToHost));
  if (taken) *taken = h->flt.taken;
  return 0;
}

int sem2d_fetch_fault_state(sem2d_solver* h, double* D, double* V, double* T, double* MU, double* Ts) {
  if (!h) return 1;
  if (!h->flt.on) return set_err(h, 4, "fetch_fault_state: no fault");
  const FaultDev& F = h->flt.dev;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  const size_t nd = (size_t)F.ndof * F.np;
  if (D) HIPCHK(h, hipMemcpy(D, F.D, nd * sizeof(double), hipMemcpyDeviceToHost));
  if (V) HIPCHK(h, hipMemcpy(V, F.V, nd * sizeof(double), hipMemcpyDeviceToHost));
  if (T) HIPCHK(h, hipMemcpy(T, F.T, 2 * F.np * sizeof(double), hipMemcpyDeviceToHost));
  if (MU) HIPCHK(h, hipMemcpy(MU, F.MU, F.np * sizeof(double), hipMemcpyDeviceToHost));
  if (Ts) HIPCHK(h, hipMemcpy(Ts, F.Tstick, 2 * F.np * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int sem2d_compute_fint(sem2d_solver* h, const double* d, double* f) {
  if (!h) return 1;
  if (!d || !f) return set_err(h, 3, "compute_fint: null array");
  const size_t n = (size_t)h->ndof * h->npoin;
  void *dd = nullptr, *ff = nullptr;
  HIPCHK(h, hipMalloc(&dd, n * rsize(h)));
  HIPCHK(h, hipMalloc(&ff, n * rsize(h)));
  int rc = write_bulk(h, dd, d, n);
  if (!rc) {
    if (hipMemsetAsync(ff, 0, n * rsize(h), h->stream) != hipSuccess) rc = set_err(h, 10, "memset failed");
  }
  if (!rc) {
    if (h->fp32) launch_force<float>(h, (const float*)dd, (float*)ff);
    else launch_force<double>(h, (const double*)dd, (double*)ff);
    if (hipGetLastError() != hipSuccess) rc = set_err(h, 10, "force kernel launch failed");
  }
  if (!rc) rc = read_bulk(h, f, ff, n);
  hipFree(dd);
  hipFree(ff);
  return rc;
}

int sem2d_get_step(sem2d_solver* h, int* it) {
  if (!h || !it) return 1;
  *it = h->it;
  return 0;
}

int sem2d_last_step_ms(sem2d_solver* h, double* ms) {
  if (!h || !ms) return 1;
  *ms = h->last_ms;
  return 0;
}

const char* sem2d_last_error(sem2d_solver* h) { return h ? h->err.c_str() : "null handle"; }

void sem2d_destroy(sem2d_solver* h) {
  if (!h) return;
  if (h->stream) hipStreamSynchronize(h->stream);
  for (void* p : h->allocs) hipFree(p);
  for (auto& b : h->abso) { (void)b; }
  if (h->src_dev) hipFree(h->src_dev);
  if (h->ampl_dev) hipFree(h->ampl_dev);
  if (h->ev0) hipEventDestroy(h->ev0);
  if (h->ev1) hipEventDestroy(h->ev1);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
}

}  // extern "C"
