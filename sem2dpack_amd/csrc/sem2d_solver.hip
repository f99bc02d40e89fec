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
#include <cstdlib>
#include <map>
#include <mutex>
#include <unordered_map>
#include <string>
#include <vector>

#include "sem2d.h"
#include "sem2d_fault.hpp"
#include "sem2d_kernels.hpp"
#include "sem2d_tile.hpp"

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
  // Stacey (P-SV): boundary elements, their edge points' global nodes [nbe][NG] and K
  // [nbe][NG][2]; per boundary node the contributions e*NG+k (element order) soff/sidx
  int nbe = 0;
  int *enode = nullptr, *soff = nullptr, *sidx = nullptr;
  double* K = nullptr;
  double* nrm = nullptr;  // [2][n] node normals of a boundary that is not flat (general path)
  // incident plane wave (general path): node coordinates [2][n], unit-amplitude Vin and
  // Tin [ndof][n], weights B [n]
  double *wxz = nullptr, *wV = nullptr, *wT = nullptr, *wB = nullptr;
  WavePar wp = {};
};

struct AbsoHost {
  std::vector<int> node;   // 0-based reference nodes
  std::vector<double> C;   // [ndof][n]
  std::vector<int> senode;  // Stacey: [nbe][NG] boundary-local node of each edge point
  std::vector<double> sK;   // [nbe][NG][2]
};

// Structured-mesh fast path (node-grid field order, fused tile kernel)
struct StructState {
  bool on = false, ready = false;
  int nx = 0, nz = 0, ezflt = 0, NX1 = 0, NZ1 = 0;
  int ROW = 0;                    // storage slots per node row (4*TX per tile column)
  std::vector<int> gmap;          // reference node (0-based) -> storage index
  std::vector<int> gridx;         // reference node (0-based) -> grid index Z*NX1 + X
  int nvl = 0;                    // vertical tile-edge lines (stored column-contiguous)
  size_t NS = 0;                  // field component stride: ROW*NZ1 + nvl*NZ1
  std::vector<int4> tiles;
  int ntiles = 0, per_xcd = 0;
  int pipe_wgs = 0;  // > 0: persistent pipelined tile kernel with this many workgroups
  // Kelvin-Voigt eta and 2.5D coefficient per element point, cell order [cell][25] (bulk
  // type), or null; either makes the tile kernel form d + eta v at the element points
  void *eta = nullptr, *b25 = nullptr;
  int pipe_threads = 512;
  int4* tiles_dev = nullptr;
  int4* grids_dev = nullptr;
  int* tlist_dev = nullptr;  // launch list of the pipelined kernel: full tiles, then the rest
  int nfull = 0, npart = 0;
  // Stacey terms: boundary elements (storage-index edge nodes, K), per absorbing entry
  // of the deferred nodes the element points it subtracts, and their values per step
  int s_nbe = 0;
  int *s_enode = nullptr, *ak_off = nullptr, *ak_idx = nullptr;
  double *s_K = nullptr, *kxd = nullptr;
  int* spec_cnt = nullptr;
  int* spec_idx = nullptr;
  void* fpart = nullptr;
  size_t FPS = 0;
  DefTable defs;
  int ndef = 0;
  int* part_off = nullptr;
  SrcEntry* dsrc = nullptr;
  double* dabso = nullptr;
  void* rmdef = nullptr;
  void* rmass0 = nullptr;
  void *coefs = nullptr, *geom = nullptr;  // cell-ordered per-element arrays
  // compact moduli once per node in the field layout ([ndof][NS], bulk type), when every
  // element sharing a node holds the same values there (COEF_NODE tile kernel)
  void* nodemod = nullptr;
  std::vector<void*> prep_allocs;  // freed when the deferred tables are rebuilt
  std::vector<int> cell_elem;      // cell (iz*nx + ix) -> element (0-based)
  std::vector<int> def_ref;        // deferred node -> reference node (0-based)
};

struct FaultHost {
  bool on = false;
  FaultDev dev;
  std::vector<void*> allocs;
  int oit, oitd, oix1, oixn, oixd, nout, nsamp, taken;
  double* rec;  // device [nsamp][6][nout]
  double* srec = nullptr;  // osides: device [nsamp][4*ndof][nout] (d, v of each side)
  double* pot = nullptr;  // device [pot_cap][2*(ndof+1)] potency rows, one per step from it = 0
  int pot_rows = 0, pot_cap = 0;
};

}  // namespace

struct AdjState;
void adj_free(AdjState* A);

struct sem2d_solver {
  int ndof, nelem, npoin, nelast, coef_mode, fp32;
  size_t fstride = 0;  // component stride of the device field arrays
  double dt, beta, gamma;
  int leapfrog = 0;  // 1: central-difference scheme (SRC/solver.f90:140-160)
  int hht = 0;       // 1: HHT-alpha (SRC/solver.f90:89-131): forces at d, v of n + alpha
  double alpha = 1.0;
  void *da = nullptr, *va = nullptr;  // the fields at n + alpha (bulk type)
  int ng = NG, np = NP;     // NGLL and GLL points per element
  double* hg = nullptr;     // h'_i(x_j) at hg[i*ng+j] (general element kernel, Stacey)
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // bulk arrays (type R = double or float, stored as raw bytes)
  void *d = nullptr, *v = nullptr, *a = nullptr, *rmass = nullptr, *coefs = nullptr, *geom = nullptr;
  int* ibool = nullptr;  // colour-sorted, 0-based
  std::vector<int> colour_start;  // ncolour+1 offsets into the sorted order
  std::vector<int> elem_order;    // sorted position -> original element (0-based)
  // Dirichlet boundaries (BC_DIRNEU 'D' components), in registration order
  struct DirBC {
    std::vector<int> node;  // 0-based reference nodes
    int comp = 0;           // bit c: component c set to zero
    bool after_fault = false;
    int* node_dev = nullptr;
  };
  std::vector<DirBC> dirs;
  // boundaries and sources
  std::vector<AbsoDev> abso;
  std::vector<AbsoHost> abso_host;
  std::vector<double> rmass_host;          // [ndof][npoin], reference order
  std::vector<int> fault_n1, fault_n2;     // 0-based reference nodes
  StructState st;
  std::vector<SrcEntry> src_host;
  std::vector<std::vector<double>> stf;
  int nsrc = 0;
  SrcEntry* src_dev = nullptr;
  double* ampl_dev = nullptr;
  int ampl_nt = 0;
  bool src_dirty = false;
  // receivers
  int nrec = 0, isamp = 1, field = 1, nsamp = 0, taken = 0;
  int* rec_node = nullptr;       // [nrec] node, or [nrec][25] element nodes when interpolated
  double* rec_w = nullptr;       // [nrec][25] interpolation weights (nullptr: receivers at nodes)
  double* seis = nullptr;
  FaultHost flt;
  int* flags = nullptr;  // [0] Newton failure
  int it = 0;
  double last_ms = 0.0;
  std::string err;
  std::vector<void*> allocs;
  AdjState* adj = nullptr;
  std::vector<int> elem_ibool;  // [nelem][25] 0-based reference nodes (interpolated receivers)
  // linear slip fault (BC_LSF): node pairs, side-1 normals [2][n], B, W1, W2 [ndof][n]
  int nlsf = 0;
  int *lsf_n1 = nullptr, *lsf_n2 = nullptr;
  double *lsf_nrm = nullptr, *lsf_B = nullptr, *lsf_W1 = nullptr, *lsf_W2 = nullptr;
  double lsf_kt = 0, lsf_kn = 0;
  int nper = 0;                 // periodic node pairs (master, slave; 0-based)
  int *per_m = nullptr, *per_s = nullptr;
  void* kv_eta = nullptr;       // Kelvin-Voigt viscosity [element in colour order][25] (bulk type)
  void* beta25 = nullptr;       // 2.5D coefficients [element in colour order][25] (bulk type)
  // progress rows (it, time, max |v|, max |d|) every itinfo steps, and the non-finite
  // watchdog (k_field_max, k_info_row)
  int itinfo = 0, info_rows = 0, info_cap = 0;
  double* info = nullptr;                 // device [info_cap][4]
  unsigned long long* info_acc = nullptr;  // device [2]
  int* info_cnt = nullptr;                 // device [2]
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

// Scratch memory of sem2d_compute_fint_noalloc: a file-scope device array, one copy per
// GPU, held until the process ends.  Host code never passes the array's name to a kernel
// or a copy (on the host it names a stand-in, not device memory); it uses the device
// address taken once per device with hipGetSymbolAddress (scratch_on_device).
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

// device address of this device's scratch array (taken at set-up, once per device, with
// the device current); nullptr with an error message on failure
void* scratch_on_device(sem2d_solver* h) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
    set_err(h, 10, "scratch: no current device");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_err(h, 10, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

// staging kernels of sem2d_compute_fint_noalloc (grid-stride)
template <typename T>
__global__ void k_fill(T* __restrict__ x, T val, size_t n) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
    x[k] = val;
}
__global__ void k_to_float(const double* __restrict__ x, float* __restrict__ y, size_t n) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
    y[k] = (float)x[k];
}
__global__ void k_to_double(const float* __restrict__ x, double* __restrict__ y, size_t n) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
    y[k] = (double)x[k];
}

// Element colouring: no two elements of one colour share a node.
// Structured boxes use the 2x2 checkerboard of cells; otherwise greedy first-fit
// in element order.
int build_colours(sem2d_solver* h, const sem2d_desc* s, std::vector<int>& colour, int& ncol) {
  const int ne = s->nelem;
  const int NP = s->ngll * s->ngll;  // points per element
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

template <typename R, int NDOF, int NELAST, int MODE, int NGL>
void launch_force_t(sem2d_solver* h, const R* d, R* f, const R* v) {
  const R* kv = v ? (const R*)h->kv_eta : nullptr;
  constexpr int EPB = fe_epb(NGL);
  for (size_t c = 0; c + 1 < h->colour_start.size(); ++c) {
    const int e0 = h->colour_start[c], ne = h->colour_start[c + 1] - e0;
    if (ne <= 0) continue;
    k_force_colour<R, NDOF, NELAST, MODE, NGL><<<(ne + EPB - 1) / EPB, 256, 0, h->stream>>>(
        e0, ne, h->ibool, (const R*)h->coefs, (const R*)h->geom, d, f, h->npoin, v, kv, (const R*)h->beta25,
        h->hg);
  }
}

template <typename R, int NDOF, int NELAST, int NGL>
void launch_force_m(sem2d_solver* h, const R* d, R* f, const R* v) {
  switch (h->coef_mode) {
    case COEF_SHARED: launch_force_t<R, NDOF, NELAST, COEF_SHARED, NGL>(h, d, f, v); break;
    case COEF_COMPACT:  // built for NGLL = 5 only (checked at creation)
      if constexpr ((NELAST == 2 || NELAST == 6) && NGL == NG)
        launch_force_t<R, NDOF, NELAST, COEF_COMPACT, NGL>(h, d, f, v);
      break;
    default: launch_force_t<R, NDOF, NELAST, COEF_PER_ELEM, NGL>(h, d, f, v);
  }
}

template <typename R, int NGL>
void launch_force_n(sem2d_solver* h, const R* d, R* f, const R* v) {
  if (h->ndof == 2) {
    if (h->nelast == 6) launch_force_m<R, 2, 6, NGL>(h, d, f, v);
    else launch_force_m<R, 2, 10, NGL>(h, d, f, v);
  } else {
    if (h->nelast == 2) launch_force_m<R, 1, 2, NGL>(h, d, f, v);
    else launch_force_m<R, 1, 3, NGL>(h, d, f, v);
  }
}

// f = -K d, or -K (d + eta v) in Kelvin-Voigt elements when v is given
template <typename R>
void launch_force(sem2d_solver* h, const R* d, R* f, const R* v = nullptr) {
  switch (h->ng) {
    case 3: launch_force_n<R, 3>(h, d, f, v); break;
    case 4: launch_force_n<R, 4>(h, d, f, v); break;
    case 6: launch_force_n<R, 6>(h, d, f, v); break;
    case 7: launch_force_n<R, 7>(h, d, f, v); break;
    case 8: launch_force_n<R, 8>(h, d, f, v); break;
    case 9: launch_force_n<R, 9>(h, d, f, v); break;
    case 10: launch_force_n<R, 10>(h, d, f, v); break;
    default: launch_force_n<R, NG>(h, d, f, v);
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
  if (h->rec_w)
    k_receivers_interp<R><<<grid_for(h->nrec, 64), 64, 0, h->stream>>>(
        h->nrec, h->ndof, h->rec_node, h->rec_w, fld, h->fstride, h->seis + (size_t)s * h->ndof * h->nrec, h->np);
  else
    k_receivers<R><<<grid_for(h->nrec, 64), 64, 0, h->stream>>>(h->nrec, h->ndof, h->rec_node, fld, h->fstride,
                                                                h->seis + (size_t)s * h->ndof * h->nrec);
  h->taken = std::max(h->taken, s + 1);
  return 0;
}

// room for `rows` potency rows (grown outside any step loop)
int info_reserve(sem2d_solver* h, int rows) {
  if (rows <= h->info_cap) return 0;
  const int cap = std::max(rows, std::max(64, 2 * h->info_cap));
  double* p = nullptr;
  HIPCHK(h, hipMalloc(&p, (size_t)cap * 4 * sizeof(double)));
  if (h->info && h->info_rows)
    HIPCHK(h, hipMemcpyAsync(p, h->info, (size_t)h->info_rows * 4 * sizeof(double), hipMemcpyDeviceToDevice,
                             h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (h->info) hipFree(h->info);
  h->info = p;
  h->info_cap = cap;
  return 0;
}

// the progress row of step it (after that step's kernels)
template <typename R>
void record_info(sem2d_solver* h, int it) {
  if (h->itinfo <= 0 || it % h->itinfo != 0) return;
  const size_t n = (size_t)h->ndof * h->fstride;
  k_field_max<R><<<std::min<size_t>((n + 255) / 256, 4096), 256, 0, h->stream>>>(
      n, (const R*)h->d, (const R*)h->v, (const R*)h->a, h->info_acc, h->info_cnt);
  k_info_row<<<1, 64, 0, h->stream>>>(it, it * h->dt, h->info_acc, h->info_cnt, h->info + 4 * (size_t)h->info_rows);
  ++h->info_rows;
}

int potency_reserve(sem2d_solver* h, int rows) {
  FaultHost& F = h->flt;
  if (rows <= F.pot_cap) return 0;
  const int cap = std::max(rows, 2 * F.pot_cap);
  const size_t nc = 2 * (h->ndof + 1);
  double* p = nullptr;
  HIPCHK(h, hipMalloc(&p, (size_t)cap * nc * sizeof(double)));
  if (F.pot && F.pot_rows)
    HIPCHK(h, hipMemcpyAsync(p, F.pot, (size_t)F.pot_rows * nc * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (F.pot) hipFree(F.pot);
  F.pot = p;
  F.pot_cap = cap;
  return 0;
}

template <typename R>
void record_potency(sem2d_solver* h) {
  FaultHost& F = h->flt;
  if (F.pot_rows >= F.pot_cap) return;  // reserved by the caller
  k_fault_potency<R><<<1, 256, 0, h->stream>>>(F.dev, (const R*)h->d, (const R*)h->v, h->fstride,
                                               F.pot + (size_t)F.pot_rows * 2 * (h->ndof + 1));
  F.pot_rows++;
}

// per-side record (osides) of sample slot s from the current fields
void record_sides(sem2d_solver* h, int s) {
  FaultHost& F = h->flt;
  double* out = F.srec + (size_t)s * 4 * h->ndof * F.nout;
  if (h->fp32)
    k_fault_sides<float><<<grid_for(F.nout, 64), 64, 0, h->stream>>>(F.dev, (const float*)h->d, (const float*)h->v,
                                                                      h->fstride, F.oix1, F.oixd, F.nout, out);
  else
    k_fault_sides<double><<<grid_for(F.nout, 64), 64, 0, h->stream>>>(
        F.dev, (const double*)h->d, (const double*)h->v, h->fstride, F.oix1, F.oixd, F.nout, out);
}

int record_fault(sem2d_solver* h) {
  FaultHost& F = h->flt;
  if (!F.on) return 0;
  if (h->fp32) record_potency<float>(h);
  else record_potency<double>(h);
  if (h->it < F.oit) return 0;
  if (F.taken < F.nsamp) {
    k_fault_record<<<grid_for(F.nout, 64), 64, 0, h->stream>>>(F.dev, F.oix1, F.oixd, F.nout,
                                                               F.rec + (size_t)F.taken * 6 * F.nout);
    if (F.srec) record_sides(h, F.taken);
    F.taken++;
  }
  F.oit += F.oitd;
  return 0;
}

// ---------------------------------------------------------------------------
// structured path set-up
// ---------------------------------------------------------------------------
// Grid map from ibool and the cell -> element table; returns false if ibool does not
// number a MESH_CART box the way the node grid does (the general path is used then).
// Storage index of grid node (X, Z): nodes on vertical tile-edge lines (X a multiple
// of 4*TX, or the right domain edge) live column-contiguous after the grid block, so
// that the column edges of the tiles are read and written with unit stride.
// The other nodes are stored row by row, each tile column's 4*TX-1 nodes in a segment
// of 4*TX slots (ROW = 4*TX * tile columns slots per row), so every segment starts on
// a cache-line boundary.
static size_t storage_index(const StructState& S, int X, int Z) {
  const int W = 4 * TX;
  int k = -1;
  if (X % W == 0) k = X / W;
  else if (X == S.NX1 - 1) k = S.nvl - 1;
  if (k >= 0) return (size_t)S.ROW * S.NZ1 + (size_t)k * S.NZ1 + Z;
  return (size_t)Z * S.ROW + (size_t)(X / W) * W + (X % W) - 1;
}

bool structured_map(sem2d_solver* h, const sem2d_desc* s) {
  StructState& S = h->st;
  const int nx = s->nx, nz = s->nz, ez = s->ezflt;
  if (nx <= 0 || nz <= 0 || !s->elem_of_cell || (long)nx * nz != s->nelem) return false;
  if (nx > 0x7fff || ez < 0 || ez >= nz) return false;
  S.nx = nx; S.nz = nz; S.ezflt = ez;
  S.NX1 = 4 * nx + 1;
  S.NZ1 = 4 * nz + 1 + (ez > 0 ? 1 : 0);
  if ((long)S.NX1 * S.NZ1 != s->npoin) return false;
  S.nvl = (nx + TX - 1) / TX + 1;
  S.ROW = 4 * TX * (S.nvl - 1);
  // component stride rounded to 4*TX slots so that every component (and every batch
  // member placed at a multiple of it) keeps the segments line-aligned
  S.NS = (size_t)S.ROW * S.NZ1 + (size_t)S.nvl * S.NZ1;
  S.NS = (S.NS + 4 * TX - 1) / (4 * TX) * (4 * TX);
  S.gmap.assign(s->npoin, -1);
  std::vector<char> used((size_t)s->npoin, 0);
  for (int iz = 0; iz < nz; ++iz) {
    const int sh = (ez > 0 && iz >= ez) ? 1 : 0;
    for (int ix = 0; ix < nx; ++ix) {
      const int e = s->elem_of_cell[(size_t)iz * nx + ix] - 1;
      if (e < 0 || e >= s->nelem) return false;
      for (int j = 0; j < NG; ++j)
        for (int i = 0; i < NG; ++i) {
          const int ref = s->ibool[(size_t)e * NP + j * NG + i] - 1;
          const int g = (4 * iz + j + sh) * S.NX1 + 4 * ix + i;
          if (S.gmap[ref] == -1) {
            if (used[g]) return false;
            S.gmap[ref] = g;
            used[g] = 1;
          } else if (S.gmap[ref] != g) {
            return false;
          }
        }
    }
  }
  for (int k = 0; k < s->npoin; ++k)
    if (S.gmap[k] < 0) return false;
  S.gridx = S.gmap;
  for (int k = 0; k < s->npoin; ++k) S.gmap[k] = (int)storage_index(S, S.gridx[k] % S.NX1, S.gridx[k] / S.NX1);
  // tiles: element-row bands split at the fault, TX x TZ tiles in each band
  S.tiles.clear();
  const int nb = ez > 0 ? 2 : 1;
  for (int b = 0; b < nb; ++b) {
    const int r0 = (b == 0) ? 0 : ez, r1 = (nb == 1) ? nz : (b == 0 ? ez : nz);
    const int sh = b;
    for (int tz0 = r0; tz0 < r1; tz0 += TZ)
      for (int tx0 = 0; tx0 < nx; tx0 += TX) {
        const int nxe = std::min(TX, nx - tx0), nze = std::min(TZ, r1 - tz0);
        S.tiles.push_back(make_int4(tx0, tz0, nxe | (nze << 8) | (nx << 16), 4 * tz0 + sh));
      }
  }
  S.ntiles = (int)S.tiles.size();
  S.per_xcd = (S.ntiles + 7) / 8;
  {
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    // measured on MI355X (2000x1000-element P-SV, full-tile variant): 512 threads
    // (20-element batches, one batch per colour) with 1 workgroup per CU for FP64 fields
    // (VGPRs allow 8 waves per CU), 2 for FP32
    int per_cu = h->fp32 ? 2 : 1;
    const char* env = getenv("SEM2D_TILE_WG_PER_CU");
    if (env) per_cu = atoi(env);
    S.pipe_threads = 512;
    const char* envt = getenv("SEM2D_TILE_THREADS");
    if (envt) S.pipe_threads = atoi(envt) == 256 ? 256 : 512;
    // the persistent kernel pays off once there are several tiles per workgroup
    S.pipe_wgs = per_cu > 0 && S.ntiles >= 4 * ncu * per_cu ? ncu * per_cu : 0;
    S.pipe_wgs = ((S.pipe_wgs + 7) / 8) * 8;
  }
  return true;
}

// Launch list of the pipelined tile kernel: the full TX x TZ tiles (its constant-shape,
// 32-bit-offset variant) first, then the others.  idx32: the field and partial-force
// arrays are below 4 GiB per component, as that variant requires; otherwise every tile
// goes to the general variant.
static std::vector<int> tile_list(const std::vector<int4>& tiles, bool idx32, int& nfull) {
  std::vector<int> list;
  list.reserve(tiles.size());
  auto full = [&](const int4& tg) { return idx32 && (tg.z & 0xffff) == (TX | (TZ << 8)); };
  for (int t = 0; t < (int)tiles.size(); ++t)
    if (full(tiles[t])) list.push_back(t);
  nfull = (int)list.size();
  for (int t = 0; t < (int)tiles.size(); ++t)
    if (!full(tiles[t])) list.push_back(t);
  return list;
}

// ngp, fps: field and partial-force component strides; ncoef: entries of the per-cell
// coefficient array
static bool fits32(size_t ngp, size_t fps, size_t ncoef, size_t rs) {
  return std::max(std::max(ngp, fps), ncoef) * rs < (size_t(1) << 32);
}

// device form of a deferred-node list (DefDev records and their side table); false when
// a count does not fit its 8 bits
bool pack_defs(const std::vector<DefNode>& in, const std::vector<int>& part_off, std::vector<DefDev>& node,
               std::vector<int2>& ext) {
  node.resize(in.size());
  ext.clear();
  for (size_t q = 0; q < in.size(); ++q) {
    const DefNode& n = in[q];
    if (n.pcnt < 0 || n.pcnt > 255 || n.scnt < 0 || n.scnt > 255 || n.acnt < 0 || n.acnt > 255) return false;
    DefDev e;
    e.g = n.g;
    e.pstart = n.pstart;
    e.cnt = n.pcnt | (n.scnt << 8) | (n.acnt << 16) | ((n.fault ? 1 : 0) << 24) | ((n.dmask & 3) << 26);
    e.x = -1;
    if (n.scnt > 0 || n.acnt > 0) {
      e.x = (int)ext.size();
      ext.push_back(make_int2(n.sstart, n.astart));
    } else if (n.pcnt == 2) {  // offsets inline
      e.cnt |= 1 << 25;
      e.pstart = part_off[n.pstart];
      e.x = part_off[n.pstart + 1];
    }
    node[q] = e;
  }
  return true;
}

template <typename T>
int prep_upload(sem2d_solver* h, T** p, const T* src, size_t n) {
  void* q = nullptr;
  HIPCHK(h, hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
  h->st.prep_allocs.push_back(q);
  *p = (T*)q;
  if (n) HIPCHK(h, hipMemcpyAsync(q, src, n * sizeof(T), hipMemcpyHostToDevice, h->stream));
  return 0;
}

int prep_upload_bulk(sem2d_solver* h, void** p, const std::vector<double>& src) {
  if (!h->fp32) return prep_upload(h, (double**)p, src.data(), src.size());
  std::vector<float> tmp(src.begin(), src.end());
  int rc = prep_upload(h, (float**)p, tmp.data(), tmp.size());
  if (rc) return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

// Host tables of the structured step of one grid, with local indices (grid node
// storage from 0, tiles from 0, cells from 0): tiles and their grid descriptors,
// special nodes, deferred nodes with their partial-force slots, sources, absorbing
// coefficients and inverse masses, and the split-fault nodes.
struct StructTables {
  std::vector<int4> tiles, grids;
  std::vector<int> spec_cnt, spec_idx;
  std::vector<DefNode> defs;
  std::vector<int> part_off;
  std::vector<SrcEntry> dsrc;
  std::vector<double> dabso;
  std::vector<double> rmdef;     // [ndof][ndef]
  std::vector<int> fault_g;
  std::vector<double> fault_rm;  // [ndof][nf]
  std::vector<unsigned char> fault_dmask;  // [npairs] Dirichlet after the fault: node 1 bits 0-1, node 2 bits 2-3
  std::vector<int> s_enode, ak_off, ak_idx;  // Stacey (see StructState)
  std::vector<double> s_K;
};

int struct_tables(sem2d_solver* h, StructTables& X) {
  StructState& S = h->st;
  const int nd = h->ndof;
  X.tiles = S.tiles;
  X.grids.assign(S.ntiles, make_int4(0, S.ROW, S.NZ1, 0));
  // special nodes: point-source nodes inside a tile patch
  std::vector<int> src_g;  // grid indices of source nodes
  for (const SrcEntry& e : h->src_host) src_g.push_back(S.gridx[e.node]);
  std::sort(src_g.begin(), src_g.end());
  src_g.erase(std::unique(src_g.begin(), src_g.end()), src_g.end());
  X.spec_cnt.assign(S.ntiles, 0);
  X.spec_idx.assign((size_t)S.ntiles * MAXSPECIAL, 0);
  // entries (g, tile, slot)
  struct Ent { int g, t, slot; };
  std::vector<Ent> ents;
  ents.reserve((size_t)S.ntiles * PERIM_MAX);
  for (int t = 0; t < S.ntiles; ++t) {
    const int4 tg = S.tiles[t];
    const int nxe = tg.z & 0xff, nze = (tg.z >> 8) & 0xff;
    const int PX = 4 * nxe + 1, PZ = 4 * nze + 1, X0 = 4 * tg.x, Z0 = tg.w;
    auto gi = [&](int Xn, int Zn) { return (int)storage_index(S, X0 + Xn, Z0 + Zn); };
    for (int Xn = 0; Xn < PX; ++Xn) ents.push_back({gi(Xn, 0), t, Xn});
    for (int Xn = 0; Xn < PX; ++Xn) ents.push_back({gi(Xn, PZ - 1), t, PX + Xn});
    for (int Zn = 1; Zn < PZ - 1; ++Zn) ents.push_back({gi(0, Zn), t, 2 * PX + Zn - 1});
    for (int Zn = 1; Zn < PZ - 1; ++Zn) ents.push_back({gi(PX - 1, Zn), t, 2 * PX + (PZ - 2) + Zn - 1});
    for (int g : src_g) {
      const int Xn = g % S.NX1 - X0, Zn = g / S.NX1 - Z0;
      if (Xn > 0 && Xn < PX - 1 && Zn > 0 && Zn < PZ - 1) {
        if (X.spec_cnt[t] >= MAXSPECIAL) return set_err(h, 5, "too many source nodes in one tile");
        X.spec_idx[(size_t)t * MAXSPECIAL + X.spec_cnt[t]] = Zn * PX + Xn;
        ents.push_back({(int)storage_index(S, X0 + Xn, Z0 + Zn), t, PERIM_MAX + X.spec_cnt[t]});
        X.spec_cnt[t]++;
      }
    }
  }
  std::sort(ents.begin(), ents.end(), [](const Ent& a, const Ent& b) {
    return a.g != b.g ? a.g < b.g : a.t < b.t;
  });
  // inverse map storage -> reference node, for sources / ABC / fault lookups
  std::vector<int> ginv(S.NS, -1);
  for (int k = 0; k < h->npoin; ++k) ginv[S.gmap[k]] = k;
  std::unordered_map<int, std::vector<int>> src_of;  // ref node -> src_host indices
  for (size_t q = 0; q < h->src_host.size(); ++q) src_of[h->src_host[q].node].push_back((int)q);
  std::unordered_map<int, std::vector<std::pair<int, int>>> abso_of;  // ref node -> (bc, k)
  for (size_t b = 0; b < h->abso_host.size(); ++b)
    for (size_t k = 0; k < h->abso_host[b].node.size(); ++k)
      abso_of[h->abso_host[b].node[k]].push_back({(int)b, (int)k});
  std::vector<char> is_fault((size_t)h->npoin, 0);
  for (int k : h->fault_n1) is_fault[k] = 1;
  for (int k : h->fault_n2) is_fault[k] = 1;
  // Dirichlet components per reference node: those applied by k_deferred (every
  // Dirichlet boundary at a node off the fault; at a fault node only those that act
  // before the fault) and those k_fault applies after the fault solve
  std::vector<unsigned char> dmask_def((size_t)h->npoin, 0), dmask_aft((size_t)h->npoin, 0);
  for (const auto& D : h->dirs)
    for (int k : D.node) {
      if (is_fault[k] && D.after_fault) dmask_aft[k] |= (unsigned char)D.comp;
      else dmask_def[k] |= (unsigned char)D.comp;
    }
  // Stacey: global list of boundary elements (in boundary order) and, per boundary and
  // boundary-local node, its element points e*NG+k in element order
  X.s_enode.clear();
  X.s_K.clear();
  std::vector<std::vector<std::vector<int>>> scon(h->abso_host.size());
  for (size_t b = 0; b < h->abso_host.size(); ++b) {
    const AbsoHost& B = h->abso_host[b];
    if (B.senode.empty()) continue;
    const int base = (int)(X.s_enode.size() / NG), nbe = (int)(B.senode.size() / NG);
    scon[b].assign(B.node.size(), {});
    for (int e = 0; e < nbe; ++e)
      for (int k = 0; k < NG; ++k) {
        const int bn = B.senode[(size_t)e * NG + k];
        X.s_enode.push_back(S.gmap[B.node[bn]]);
        scon[b][bn].push_back((base + e) * NG + k);
      }
    X.s_K.insert(X.s_K.end(), B.sK.begin(), B.sK.end());
  }
  X.ak_off.assign(1, 0);
  X.ak_idx.clear();

  X.defs.clear();
  X.part_off.clear();
  X.dsrc.clear();
  X.dabso.clear();
  std::vector<int> def_ref;
  for (size_t i = 0; i < ents.size();) {
    size_t j = i;
    DefNode n{};
    n.g = ents[i].g;
    n.pstart = (int)X.part_off.size();
    while (j < ents.size() && ents[j].g == n.g) {
      X.part_off.push_back(ents[j].t * SLOTS + ents[j].slot);
      ++j;
    }
    n.pcnt = (int)X.part_off.size() - n.pstart;
    const int ref = ginv[n.g];
    n.sstart = (int)X.dsrc.size();
    auto its = src_of.find(ref);
    if (its != src_of.end())
      for (int q : its->second) X.dsrc.push_back(h->src_host[q]);
    n.scnt = (int)X.dsrc.size() - n.sstart;
    n.astart = (int)(X.dabso.size() / nd);
    auto ita = abso_of.find(ref);
    if (ita != abso_of.end())
      for (auto& bk : ita->second) {
        const AbsoHost& B = h->abso_host[bk.first];
        const size_t nb = B.node.size();
        for (int c = 0; c < nd; ++c) X.dabso.push_back(B.C[(size_t)c * nb + bk.second]);
        if (!scon[bk.first].empty())
          for (int t : scon[bk.first][bk.second]) X.ak_idx.push_back(t);
        X.ak_off.push_back((int)X.ak_idx.size());
      }
    n.acnt = (int)(X.dabso.size() / nd) - n.astart;
    n.fault = is_fault[ref];
    n.dmask = dmask_def[ref];
    X.defs.push_back(n);
    def_ref.push_back(ref);
    i = j;
  }
  const size_t ndef = X.defs.size();
  {  // every Dirichlet node is a deferred node (tile perimeters hold the box's sides)
    size_t ndir_def = 0, ndir = 0;
    for (size_t q = 0; q < X.defs.size(); ++q) ndir_def += X.defs[q].dmask || dmask_aft[def_ref[q]] ? 1 : 0;
    for (size_t k = 0; k < dmask_def.size(); ++k) ndir += dmask_def[k] || dmask_aft[k] ? 1 : 0;
    if (ndir_def != ndir)
      return set_err(h, 5, "structured grid: %zu of the %zu Dirichlet nodes are not on tile edges", ndir_def, ndir);
  }
  X.fault_dmask.assign(h->fault_n1.size(), 0);
  for (size_t k = 0; k < h->fault_n1.size(); ++k) {
    X.fault_dmask[k] = dmask_aft[h->fault_n1[k]];
    if (!h->fault_n2.empty()) X.fault_dmask[k] |= (unsigned char)(dmask_aft[h->fault_n2[k]] << 2);
  }
  // every split-fault node must be a deferred node: k_deferred leaves those nodes
  // uncorrected and k_fault finishes both nodes of each pair; a fault node inside a
  // tile would be corrected by the tile kernel and again by k_fault
  {
    size_t nfault_def = 0;
    for (const DefNode& n : X.defs) nfault_def += n.fault ? 1 : 0;
    if (nfault_def != h->fault_n1.size() + h->fault_n2.size())
      return set_err(h, 5, "structured grid: %zu of the %zu split-fault nodes lie on tile edges; the fault "
                     "must follow a row of element edges (ezflt) for the structured path",
                     nfault_def, h->fault_n1.size() + h->fault_n2.size());
  }
  X.rmdef.assign((size_t)nd * ndef, 0.0);
  for (size_t q = 0; q < ndef; ++q)
    for (int c = 0; c < nd; ++c) X.rmdef[(size_t)c * ndef + q] = h->rmass_host[(size_t)c * h->npoin + def_ref[q]];
  // split-fault nodes for the post-fault corrector
  X.fault_g.clear();
  for (int k : h->fault_n1) X.fault_g.push_back(S.gmap[k]);
  for (int k : h->fault_n2) X.fault_g.push_back(S.gmap[k]);
  const size_t nf = X.fault_g.size();
  X.fault_rm.assign((size_t)nd * nf, 0.0);
  for (size_t q = 0; q < nf; ++q) {
    const int ref = q < h->fault_n1.size() ? h->fault_n1[q] : h->fault_n2[q - h->fault_n1.size()];
    for (int c = 0; c < nd; ++c) X.fault_rm[(size_t)c * nf + q] = h->rmass_host[(size_t)c * h->npoin + ref];
  }
  return 0;
}

// Upload the structured tables of the solver's own grid.
// inverse masses of the split-fault nodes in pair order: r1 (node 1) and r2 (node 2),
// [ndof][npairs] each, from the table's fault_rm ([ndof][node 1 of every pair, then node 2])
// (np pairs; a one-sided fault has no node 2: r2 stays zero)
void split_fault_rm(const StructTables& X, int nd, size_t np, std::vector<double>& r1, std::vector<double>& r2) {
  const size_t nf = X.fault_g.size();
  r1.assign((size_t)nd * np, 0.0);
  r2.assign((size_t)nd * np, 0.0);
  for (int c = 0; c < nd; ++c)
    for (size_t k = 0; k < np; ++k) {
      r1[(size_t)c * np + k] = X.fault_rm[(size_t)c * nf + k];
      if (nf == 2 * np) r2[(size_t)c * np + k] = X.fault_rm[(size_t)c * nf + np + k];
    }
}

int prepare_structured(sem2d_solver* h) {
  StructState& S = h->st;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (void* q : S.prep_allocs) hipFree(q);
  S.prep_allocs.clear();
  StructTables X;
  int rc = struct_tables(h, X);
  if (rc) return rc;
  S.ndef = (int)X.defs.size();
  S.def_ref.resize(S.ndef);
  for (int q = 0; q < S.ndef; ++q) {
    // storage index -> reference node (inverse of gmap)
    S.def_ref[q] = X.defs[q].g;
  }
  {
    std::vector<int> ginv(S.NS, -1);
    for (int k = 0; k < h->npoin; ++k) ginv[S.gmap[k]] = k;
    for (int q = 0; q < S.ndef; ++q) S.def_ref[q] = ginv[S.def_ref[q]];
  }
  if ((rc = prep_upload(h, &S.tiles_dev, X.tiles.data(), X.tiles.size()))) return rc;
  if ((rc = prep_upload(h, &S.grids_dev, X.grids.data(), X.grids.size()))) return rc;
  if ((rc = prep_upload(h, &S.spec_cnt, X.spec_cnt.data(), X.spec_cnt.size()))) return rc;
  if ((rc = prep_upload(h, &S.spec_idx, X.spec_idx.data(), X.spec_idx.size()))) return rc;
  {
    std::vector<DefDev> dn;
    std::vector<int2> dx;
    if (!pack_defs(X.defs, X.part_off, dn, dx))
      return set_err(h, 5, "prepare: more than 255 entries at a deferred node");
    DefDev* pn = nullptr;
    int2* px = nullptr;
    if ((rc = prep_upload(h, &pn, dn.data(), dn.size()))) return rc;
    if ((rc = prep_upload(h, &px, dx.data(), dx.size()))) return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));  // the packed copies are local
    S.defs.node = pn;
    S.defs.ext = px;
  }
  if ((rc = prep_upload(h, &S.part_off, X.part_off.data(), X.part_off.size()))) return rc;
  if ((rc = prep_upload(h, &S.dsrc, X.dsrc.data(), X.dsrc.size()))) return rc;
  if ((rc = prep_upload(h, &S.dabso, X.dabso.data(), X.dabso.size()))) return rc;
  if ((rc = prep_upload_bulk(h, &S.rmdef, X.rmdef))) return rc;
  S.FPS = (size_t)S.ntiles * SLOTS;
  {
    // per-cell arrays: the element coefficients and the element-point terms (if any)
    size_t ncell = h->coef_mode == COEF_SHARED ? (size_t)h->nelast * NP
                                               : (size_t)h->nelem * std::max(h->ndof, h->nelast) * NP;
    if (S.eta || S.b25) ncell = std::max(ncell, (size_t)h->nelem * NP);
    const std::vector<int> list = tile_list(X.tiles, fits32(h->fstride, S.FPS, ncell, rsize(h)), S.nfull);
    S.npart = S.ntiles - S.nfull;
    if ((rc = prep_upload(h, &S.tlist_dev, list.data(), list.size()))) return rc;
  }
  void* fp = nullptr;
  HIPCHK(h, hipMalloc(&fp, std::max<size_t>(S.FPS * h->ndof, 1) * rsize(h)));
  S.prep_allocs.push_back(fp);
  S.fpart = fp;
  S.s_nbe = (int)(X.s_enode.size() / NG);
  if (S.s_nbe > 0) {
    if ((rc = prep_upload(h, &S.s_enode, X.s_enode.data(), X.s_enode.size()))) return rc;
    if ((rc = prep_upload(h, &S.s_K, X.s_K.data(), X.s_K.size()))) return rc;
    if ((rc = prep_upload(h, &S.ak_off, X.ak_off.data(), X.ak_off.size()))) return rc;
    if ((rc = prep_upload(h, &S.ak_idx, X.ak_idx.data(), X.ak_idx.size()))) return rc;
    void* q = nullptr;
    HIPCHK(h, hipMalloc(&q, X.s_enode.size() * 2 * sizeof(double)));
    S.prep_allocs.push_back(q);
    S.kxd = (double*)q;
  }
  if (h->flt.on) {  // per-pair inverse masses for the corrector inside k_fault
    std::vector<double> r1, r2;
    split_fault_rm(X, h->ndof, h->fault_n1.size(), r1, r2);
    double *p1 = nullptr, *p2 = nullptr;
    if ((rc = prep_upload(h, &p1, r1.data(), r1.size()))) return rc;
    if ((rc = prep_upload(h, &p2, r2.data(), r2.size()))) return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));  // the vectors are local
    h->flt.dev.rmn1 = p1;
    h->flt.dev.rmn2 = p2;
    h->flt.dev.dmask = nullptr;
    bool any = false;
    for (unsigned char m : X.fault_dmask) any |= m != 0;
    if (any) {
      unsigned char* pm = nullptr;
      if ((rc = prep_upload(h, &pm, X.fault_dmask.data(), X.fault_dmask.size()))) return rc;
      HIPCHK(h, hipStreamSynchronize(h->stream));
      h->flt.dev.dmask = pm;
    }
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  S.ready = true;
  return 0;
}

// Everything one structured step launches on: the device tables of one grid (a
// solver) or of many grids sharing the field arrays (a batch).
struct StepDev {
  int ndof = 0, nelast = 0, coef_mode = 0;
  size_t fstride = 0;
  void *d = nullptr, *v = nullptr, *a = nullptr;
  TileDev T{};
  const int* tlist = nullptr;  // pipelined kernel: tlist[0 .. nfull) full tiles, then npart others
  int nfull = 0, npart = 0;
  int s_nbe = 0;  // Stacey boundary elements (k_stacey_pre), their values passed to k_deferred
  const int* s_enode = nullptr;
  const double* s_K = nullptr;
  StaceyRefs SR{};
  int pipe_wgs = 0, pipe_threads = 512;
  const void *rmass0 = nullptr, *coefs = nullptr, *geom = nullptr;
  const void* nodemod = nullptr;  // per-node moduli (COEF_NODE), or null
  const void *eta = nullptr, *b25 = nullptr;  // element-point terms (Kelvin-Voigt, 2.5D), or null
  void* fpart = nullptr;
  int ndef = 0;
  DefTable defs;
  const int* part_off = nullptr;
  const SrcEntry* dsrc = nullptr;
  const double* dabso = nullptr;
  const void* rmdef = nullptr;
  const double* ampl = nullptr;
  int ampl_nt = 0;
  bool fault_on = false;
  FaultDev F{};
  const int* itp = nullptr;  // device step counter (graph replay) or null: step index by value
  hipStream_t stream = nullptr;
  // out-of-place step (adjoint replay): read the state here, write it to d, v, a
  const void *din = nullptr, *vin = nullptr, *ain = nullptr;
};

StepDev step_dev(sem2d_solver* h) {
  StructState& S = h->st;
  StepDev X;
  X.ndof = h->ndof;
  X.nelast = h->nelast;
  X.coef_mode = h->coef_mode;
  X.fstride = h->fstride;
  X.d = h->d;
  X.v = h->v;
  X.a = h->a;
  X.T.ntiles = S.ntiles;
  X.T.per_xcd = S.per_xcd;
  X.T.tiles = S.tiles_dev;
  X.T.grids = S.grids_dev;
  X.tlist = S.tlist_dev;
  X.nfull = S.nfull;
  X.npart = S.npart;
  X.s_nbe = S.s_nbe;
  X.s_enode = S.s_enode;
  X.s_K = S.s_K;
  if (S.s_nbe > 0) X.SR = StaceyRefs{S.ak_off, S.ak_idx, S.kxd};
  X.T.spec_cnt = S.spec_cnt;
  X.T.spec_idx = S.spec_idx;
  X.T.NGp = h->fstride;
  X.T.FPS = S.FPS;
  X.pipe_wgs = S.pipe_wgs;
  X.pipe_threads = S.pipe_threads;
  X.rmass0 = S.rmass0;
  X.coefs = S.coefs;
  X.nodemod = S.nodemod;
  X.eta = S.eta;
  X.b25 = S.b25;
  X.geom = S.geom;
  X.fpart = S.fpart;
  X.ndef = S.ndef;
  X.defs = S.defs;
  X.part_off = S.part_off;
  X.dsrc = S.dsrc;
  X.dabso = S.dabso;
  X.rmdef = S.rmdef;
  X.ampl = h->ampl_dev;
  X.ampl_nt = h->ampl_nt;
  X.fault_on = h->flt.on;
  if (h->flt.on) X.F = h->flt.dev;
  X.stream = h->stream;
  return X;
}

// Pipelined tile kernel over n listed tiles (FULL: all of them full tiles; EP: with the
// element-point terms, 512 threads)
template <typename R, int NDOF, int NELAST, int MODE, bool OOP, bool FULL, bool EP = false>
void launch_pipe(const StepDev& X, const int* list, int n, R dt, R c1, R c2, R c3, R c4, const R* di,
                 const R* vi, const R* ai) {
  if (n <= 0) return;
  const int wgs = std::min(X.pipe_wgs, (n + 7) / 8 * 8);
  const R* co = (const R*)(MODE == COEF_NODE ? X.nodemod : X.coefs);
  if constexpr (EP) {
    k_tile_pipe<R, NDOF, NELAST, MODE, 512, 20, OOP, FULL, true><<<wgs, 512, 0, X.stream>>>(
        X.T, list, n, X.T.tiles, X.T.grids, X.T.spec_cnt, (R*)X.d, (R*)X.v, (R*)X.a, (const R*)X.rmass0, co,
        (const R*)X.geom, (R*)X.fpart, dt, c1, c2, c3, c4, di, vi, ai, (const R*)X.eta, (const R*)X.b25);
    return;
  }
  if (X.pipe_threads == 256)
    k_tile_pipe<R, NDOF, NELAST, MODE, 256, 10, OOP, FULL><<<wgs, 256, 0, X.stream>>>(
        X.T, list, n, X.T.tiles, X.T.grids, X.T.spec_cnt, (R*)X.d, (R*)X.v, (R*)X.a, (const R*)X.rmass0, co,
        (const R*)X.geom, (R*)X.fpart, dt, c1, c2, c3, c4, di, vi, ai);
  else
    k_tile_pipe<R, NDOF, NELAST, MODE, 512, 20, OOP, FULL><<<wgs, 512, 0, X.stream>>>(
        X.T, list, n, X.T.tiles, X.T.grids, X.T.spec_cnt, (R*)X.d, (R*)X.v, (R*)X.a, (const R*)X.rmass0, co,
        (const R*)X.geom, (R*)X.fpart, dt, c1, c2, c3, c4, di, vi, ai);
}

// both launches (full tiles, then edge tiles) of the pipelined kernel; compact moduli
// are read once per node when the grid has them in that form
template <typename R, int NDOF, int NELAST, int MODE, bool OOP, bool EP = false>
void launch_pipe2(const StepDev& X, R dt, R c1, R c2, R c3, R c4, const R* di, const R* vi, const R* ai) {
  // tiles that all fit in one wave of workgroups (small grids, e.g. 100 x 100 elements):
  // a single launch of the general variant, whose duration is one tile's, instead of two
  // launches in series
  const bool one = X.nfull + X.npart <= X.pipe_wgs;
  if constexpr (MODE == COEF_COMPACT && (NELAST == 6 || NELAST == 2)) {
    if (X.nodemod) {
      if (one) {
        launch_pipe<R, NDOF, NELAST, COEF_NODE, OOP, false, EP>(X, X.tlist, X.nfull + X.npart, dt, c1, c2, c3, c4,
                                                                di, vi, ai);
        return;
      }
      launch_pipe<R, NDOF, NELAST, COEF_NODE, OOP, true, EP>(X, X.tlist, X.nfull, dt, c1, c2, c3, c4, di, vi, ai);
      launch_pipe<R, NDOF, NELAST, COEF_NODE, OOP, false, EP>(X, X.tlist + X.nfull, X.npart, dt, c1, c2, c3, c4,
                                                              di, vi, ai);
      return;
    }
  }
  if (one) {
    launch_pipe<R, NDOF, NELAST, MODE, OOP, false, EP>(X, X.tlist, X.nfull + X.npart, dt, c1, c2, c3, c4, di, vi,
                                                       ai);
    return;
  }
  launch_pipe<R, NDOF, NELAST, MODE, OOP, true, EP>(X, X.tlist, X.nfull, dt, c1, c2, c3, c4, di, vi, ai);
  launch_pipe<R, NDOF, NELAST, MODE, OOP, false, EP>(X, X.tlist + X.nfull, X.npart, dt, c1, c2, c3, c4, di, vi,
                                                     ai);
}

template <typename R, int NDOF, int NELAST, int MODE>
void launch_tile_t(const StepDev& X, R dt, R c1, R c2, R c3, R c4) {
  if constexpr (MODE == COEF_COMPACT && NDOF == 2 && NELAST == 6) {
    if (X.din) {  // out of place (only the adjoint replay, P-SV compact)
      const R *di = (const R*)X.din, *vi = (const R*)X.vin, *ai = (const R*)X.ain;
      if (X.pipe_wgs > 0) {
        launch_pipe2<R, NDOF, NELAST, MODE, true>(X, dt, c1, c2, c3, c4, di, vi, ai);
      } else {
        k_tile<R, NDOF, NELAST, MODE, true><<<8 * X.T.per_xcd, TILE_THREADS, 0, X.stream>>>(
            X.T, (R*)X.d, (R*)X.v, (R*)X.a, (const R*)X.rmass0, (const R*)X.coefs, (const R*)X.geom,
            (R*)X.fpart, dt, c1, c2, c3, c4, di, vi, ai);
      }
      return;
    }
  }
  if constexpr ((NDOF == 2 && NELAST == 6) || (NDOF == 1 && NELAST == 2)) {
    if (X.eta || X.b25) {  // Kelvin-Voigt or 2.5D: pipelined kernel with element-point terms
      launch_pipe2<R, NDOF, NELAST, MODE, false, true>(X, dt, c1, c2, c3, c4, nullptr, nullptr, nullptr);
      return;
    }
  }
  if (X.pipe_wgs > 0) {
    // persistent pipelined kernel: up to pipe_wgs workgroups, each on a run of tiles;
    // the full tiles and the others (grid edges) in two launches
    launch_pipe2<R, NDOF, NELAST, MODE, false>(X, dt, c1, c2, c3, c4, nullptr, nullptr, nullptr);
  } else {
    k_tile<R, NDOF, NELAST, MODE><<<8 * X.T.per_xcd, TILE_THREADS, 0, X.stream>>>(
        X.T, (R*)X.d, (R*)X.v, (R*)X.a, (const R*)X.rmass0, (const R*)X.coefs, (const R*)X.geom, (R*)X.fpart,
        dt, c1, c2, c3, c4);
  }
}

template <typename R, int NDOF, int NELAST>
void launch_tile_m(const StepDev& X, R dt, R c1, R c2, R c3, R c4) {
  switch (X.coef_mode) {
    case COEF_SHARED: launch_tile_t<R, NDOF, NELAST, COEF_SHARED>(X, dt, c1, c2, c3, c4); break;
    case COEF_COMPACT:
      if constexpr (NELAST == 2 || NELAST == 6) launch_tile_t<R, NDOF, NELAST, COEF_COMPACT>(X, dt, c1, c2, c3, c4);
      break;
    default: launch_tile_t<R, NDOF, NELAST, COEF_PER_ELEM>(X, dt, c1, c2, c3, c4);
  }
}

template <typename R>
void launch_structured_step(const StepDev& X, int it, R dt, R c1, R c2, R c3, R c4) {
  if (X.ndof == 2) {
    if (X.nelast == 6) launch_tile_m<R, 2, 6>(X, dt, c1, c2, c3, c4);
    else launch_tile_m<R, 2, 10>(X, dt, c1, c2, c3, c4);
  } else {
    if (X.nelast == 2) launch_tile_m<R, 1, 2>(X, dt, c1, c2, c3, c4);
    else launch_tile_m<R, 1, 3>(X, dt, c1, c2, c3, c4);
  }
  const int gd = (X.ndef + 255) / 256;
  if (X.s_nbe > 0)  // Stacey terms from the previous step's state (boundary nodes are deferred)
    k_stacey_pre<R><<<(X.s_nbe * NG + 127) / 128, 128, 0, X.stream>>>(
        X.s_nbe, X.s_enode, X.s_K, (const R*)(X.din ? X.din : X.d), (const R*)(X.din && X.vin ? X.vin : X.v),
        (const R*)(X.din && X.ain ? X.ain : X.a), X.fstride, dt, c1, (double*)X.SR.kxd);
  if (X.ndof == 2 && X.din)
    k_deferred<R, 2, true><<<gd, 256, 0, X.stream>>>(X.ndef, X.defs, X.part_off, (const R*)X.fpart, X.T.FPS,
                                                      X.dsrc, X.ampl, X.ampl_nt, it, X.itp, X.dabso,
                                                      (const R*)X.rmdef, (R*)X.d, (R*)X.v, (R*)X.a, X.fstride, dt,
                                                      c1, c2, c3, c4, (const R*)X.din, (const R*)X.vin,
                                                      (const R*)X.ain, X.SR);
  else if (X.ndof == 2)
    k_deferred<R, 2><<<gd, 256, 0, X.stream>>>(X.ndef, X.defs, X.part_off, (const R*)X.fpart, X.T.FPS, X.dsrc,
                                                X.ampl, X.ampl_nt, it, X.itp, X.dabso, (const R*)X.rmdef, (R*)X.d,
                                                (R*)X.v, (R*)X.a, X.fstride, dt, c1, c2, c3, c4, nullptr, nullptr,
                                                nullptr, X.SR);
  else
    k_deferred<R, 1><<<gd, 256, 0, X.stream>>>(X.ndef, X.defs, X.part_off, (const R*)X.fpart, X.T.FPS, X.dsrc,
                                                X.ampl, X.ampl_nt, it, X.itp, X.dabso, (const R*)X.rmdef, (R*)X.d,
                                                (R*)X.v, (R*)X.a, X.fstride, dt, c1, c2, c3, c4);
  if (X.fault_on) {
    k_fault<R, true><<<grid_for(X.F.np, 64), 64, 0, X.stream>>>(X.F, (R*)X.d, (R*)X.v, (R*)X.a, X.fstride, c3, c4,
                                                                (double)it * X.F.dt);
  }
}

// Predictor / corrector coefficients of one step: dp = d + dt v + c1 a, vp = v + c2 a,
// a' = f / M, v' = vp + c3 a', d' = dp + c4 a'.  Newmark (SRC/solver.f90:58-82):
// c1 = (1/2 - beta) dt^2, c2 = (1 - gamma) dt, c3 = gamma dt, c4 = beta dt^2.  Leapfrog
// (SRC/solver.f90:151-158: d = d + dt v_mid, v_mid = v_mid + dt a) is c1 = c2 = c4 = 0,
// c3 = dt.
template <typename R>
struct StepCoefs {
  R dt, c1, c2, c3, c4;
};
template <typename R>
StepCoefs<R> step_coefs(double dt, double beta, double gamma, int leapfrog) {
  if (leapfrog) return {(R)dt, (R)0, (R)0, (R)dt, (R)0};
  return {(R)dt, (R)((0.5 - beta) * dt * dt), (R)((1.0 - gamma) * dt), (R)(gamma * dt), (R)(beta * dt * dt)};
}

// Periodic boundaries (BC_PERIO_apply, SRC/bc_periodic.f90:78-90): the force of master
// node m[k] becomes f(m) + f(s), slave node s[k] takes it.  The pairs share no node.
template <typename R>
__global__ void k_periodic(int n, const int* __restrict__ m, const int* __restrict__ s, R* __restrict__ f,
                           int npoin, int ndof) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  for (int c = 0; c < ndof; ++c) {
    const size_t o = (size_t)c * npoin;
    const R x = f[o + m[k]] + f[o + s[k]];
    f[o + m[k]] = x;
    f[o + s[k]] = x;
  }
}

template <typename R>
int do_steps(sem2d_solver* h, int nsteps) {
  int rc = sync_sources(h);
  if (rc) return rc;
  const size_t n = (size_t)h->ndof * h->npoin;
  R* d = (R*)h->d;
  R* v = (R*)h->v;
  R* a = (R*)h->a;
  const StepCoefs<R> K = step_coefs<R>(h->dt, h->beta, h->gamma, h->leapfrog);
  const R dt = K.dt, c1 = K.c1, c2 = K.c2, c3 = K.c3, c4 = K.c4;
  const int gb = grid_for(n, 256);
  if (h->st.on && !h->st.ready && (rc = prepare_structured(h))) return rc;
  if (h->flt.on && (rc = potency_reserve(h, h->flt.pot_rows + nsteps))) return rc;
  if (h->itinfo > 0 && (rc = info_reserve(h, h->info_rows + (h->it + nsteps) / h->itinfo - h->it / h->itinfo)))
    return rc;
  StepDev X;
  if (h->st.on) X = step_dev(h);
  HIPCHK(h, hipEventRecord(h->ev0, h->stream));
  for (int s = 0; s < nsteps; ++s) {
    const int it = h->it + 1;
    if (h->st.on) {
      launch_structured_step<R>(X, it, dt, c1, c2, c3, c4);
      h->it = it;
      record_receivers<R>(h);
      record_fault(h);
      record_info<R>(h, it);
      continue;
    }
    // HHT-alpha: the forces, absorbing terms and sources see the fields at n + alpha
    R* df = d;
    R* vf = v;
    if (h->hht) {
      df = (R*)h->da;
      vf = (R*)h->va;
      k_predict_hht<R><<<gb, 256, 0, h->stream>>>(n, d, v, a, df, vf, dt, c1, c2, (R)h->alpha,
                                                  (R)(1.0 - h->alpha));
    } else {
      k_predict<R><<<gb, 256, 0, h->stream>>>(n, d, v, a, dt, c1, c2);
    }
    launch_force<R>(h, df, a, h->kv_eta ? vf : nullptr);
    if (!h->src_host.empty())
      k_sources<R><<<1, 64, 0, h->stream>>>((int)h->src_host.size(), h->src_dev, h->ampl_dev, h->ampl_nt, it, a,
                                           h->npoin);
    // periodic boundaries first among the boundary conditions (SRC/bc_gen.f90:273-275)
    if (h->nper)
      k_periodic<R><<<grid_for(h->nper, 128), 128, 0, h->stream>>>(h->nper, h->per_m, h->per_s, a, h->npoin,
                                                                  h->ndof);
    // the boundary conditions' time: it*dt, or t_alpha = it*dt + (alpha-1) dt for HHT-alpha
    // (SRC/main.f90:53, SRC/solver.f90:114-119)
    const double tbc = h->hht ? (double)it * h->dt + (h->alpha - 1.0) * h->dt : (double)it * h->dt;
    for (auto& b : h->abso) {
      if (b.wxz)
        k_abso_wave<R><<<grid_for(b.n, 128), 128, 0, h->stream>>>(b.n, h->ndof, b.node, b.C, b.wxz, b.wV, b.wT,
                                                                  b.wB, b.wp, tbc, vf, a, h->npoin);
      else if (b.nrm)
        k_abso_normal<R><<<grid_for(b.n, 128), 128, 0, h->stream>>>(b.n, b.node, b.C, b.nrm, vf, a, h->npoin);
      else
        k_abso<R><<<grid_for(b.n, 128), 128, 0, h->stream>>>(b.n, h->ndof, b.node, b.C, vf, a, h->npoin);
      if (b.nbe)
        k_stacey<R><<<grid_for(b.n, 128), 128, 0, h->stream>>>(b.n, b.node, b.soff, b.sidx, b.enode, b.K, df, a,
                                                                h->npoin, h->ng, h->hg);
    }
    // the other boundary conditions in their order (SRC/bc_gen.f90:273-281)
    for (auto& D : h->dirs)
      if (!D.after_fault)
        k_dirichlet<R><<<grid_for(D.node.size(), 128), 128, 0, h->stream>>>((int)D.node.size(), D.node_dev, D.comp,
                                                                           h->ndof, a, h->npoin);
    if (h->nlsf)
      k_lsf<R><<<grid_for(h->nlsf, 128), 128, 0, h->stream>>>(h->nlsf, h->ndof, h->lsf_n1, h->lsf_n2, h->lsf_nrm,
                                                             h->lsf_B, h->lsf_W1, h->lsf_W2, h->lsf_kt, h->lsf_kn,
                                                             df, a, h->npoin);
    if (h->flt.on)
      k_fault<R><<<grid_for(h->flt.dev.np, 64), 64, 0, h->stream>>>(h->flt.dev, d, v, a, h->npoin, R(0), R(0),
                                                                    tbc);
    for (auto& D : h->dirs)
      if (D.after_fault)
        k_dirichlet<R><<<grid_for(D.node.size(), 128), 128, 0, h->stream>>>((int)D.node.size(), D.node_dev, D.comp,
                                                                           h->ndof, a, h->npoin);
    k_correct<R><<<gb, 256, 0, h->stream>>>(n, d, v, a, (const R*)h->rmass, c3, c4);
    h->it = it;
    record_receivers<R>(h);
    record_fault(h);
    record_info<R>(h, it);
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
  if (h->itinfo > 0) {
    int cnt[2] = {0, 0};
    HIPCHK(h, hipMemcpy(cnt, h->info_cnt, sizeof cnt, hipMemcpyDeviceToHost));
    if (cnt[1]) return set_err(h, 21, "non-finite field values at step %d", cnt[1]);
  }
  return 0;
}

}  // namespace

extern "C" {

int sem2d_create(const sem2d_desc* s, sem2d_solver** out) {
  if (!s || !out) return 1;
  *out = nullptr;
  if (s->ngll < 3 || s->ngll > NGMAX) return 2;
  if (!(s->ndof == 1 || s->ndof == 2)) return 2;
  if (s->ndof == 2 && s->nelast != 6 && s->nelast != 10) return 2;
  if (s->ndof == 1 && s->nelast != 2 && s->nelast != 3) return 2;
  if (s->nelem <= 0 || s->npoin <= 0 || !s->ibool || !s->hprime || !s->rmass) return 2;
  if (s->coef_mode < 0 || s->coef_mode > 2) return 2;
  if (s->coef_mode == COEF_COMPACT &&
      (!s->lam_mu || !s->geom || !s->wgll || (s->nelast != 2 && s->nelast != 6) || s->ngll != NG))
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
  h->ng = s->ngll;
  h->np = s->ngll * s->ngll;
  const int NG = h->ng, NP = h->np;  // this problem's NGLL from here on
  for (size_t k = 0; k < (size_t)s->nelem * NP; ++k)
    if (s->ibool[k] < 1 || s->ibool[k] > s->npoin) return set_err(h, 3, "ibool entry %zu out of range", k);
  h->elem_ibool.resize((size_t)s->nelem * NP);
  for (size_t k = 0; k < (size_t)s->nelem * NP; ++k) h->elem_ibool[k] = s->ibool[k] - 1;
  HIPCHK(h, hipStreamCreate(&h->stream));
  HIPCHK(h, hipEventCreate(&h->ev0));
  HIPCHK(h, hipEventCreate(&h->ev1));
  if (!scratch_on_device(h)) return 10;  // set-up of sem2d_compute_fint_noalloc's scratch
  std::vector<double> Hc((size_t)NP);
  for (int i = 0; i < NG; ++i)
    for (int j = 0; j < NG; ++j) Hc[i * NG + j] = s->hprime[j * NG + i];
  int rc = upload(h, &h->hg, Hc.data(), Hc.size());
  if (rc) return rc;
  // the NGLL = 5 kernels read H from constant memory (the same values for every problem)
  if (NG == sem2d::NG)
    HIPCHK(h, hipMemcpyToSymbol(HIP_SYMBOL(c_H), Hc.data(), NP * sizeof(double), 0, hipMemcpyHostToDevice));
  if (s->coef_mode == COEF_COMPACT) {
    double W2[sem2d::NP];
    for (int j = 0; j < NG; ++j)
      for (int i = 0; i < NG; ++i) W2[j * NG + i] = s->wgll[i] * s->wgll[j];
    HIPCHK(h, hipMemcpyToSymbol(HIP_SYMBOL(c_W2), W2, sizeof W2, 0, hipMemcpyHostToDevice));
  }

  std::vector<int> colour;
  int ncol = 0;
  rc = build_colours(h, s, colour, ncol);
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
  h->rmass_host.assign(s->rmass, s->rmass + nf);
  h->fstride = (size_t)h->npoin;
  // tile kernels: NGLL 5 and flat-element coefficients (nelast 6 or 2)
  if (NG == sem2d::NG && (h->nelast == 6 || h->nelast == 2) && structured_map(h, s)) {
    StructState& S = h->st;
    S.on = true;
    S.cell_elem.resize(s->nelem);
    for (int cell = 0; cell < s->nelem; ++cell) S.cell_elem[cell] = s->elem_of_cell[cell] - 1;
    h->fstride = S.NS;
    std::vector<double> rm0(S.NS, 0.0);
    for (int k = 0; k < h->npoin; ++k) rm0[S.gmap[k]] = s->rmass[k];
    if ((rc = upload_bulk(h, &S.rmass0, rm0.data(), rm0.size()))) return rc;
    if (h->coef_mode == COEF_SHARED) {
      S.coefs = h->coefs;
    } else {
      // per-element arrays in cell order (iz*nx + ix)
      const size_t per = h->coef_mode == COEF_COMPACT ? (size_t)h->ndof * NP : (size_t)h->nelast * NP;
      const double* src = h->coef_mode == COEF_COMPACT ? s->lam_mu : s->coefs;
      std::vector<double> co((size_t)s->nelem * per);
      for (int cell = 0; cell < s->nelem; ++cell)
        std::memcpy(&co[(size_t)cell * per], src + (size_t)(s->elem_of_cell[cell] - 1) * per, per * sizeof(double));
      if ((rc = upload_bulk(h, &S.coefs, co.data(), co.size()))) return rc;
      if (h->coef_mode == COEF_COMPACT) {
        std::vector<double> ge((size_t)s->nelem * 3);
        for (int cell = 0; cell < s->nelem; ++cell)
          for (int q = 0; q < 3; ++q) ge[(size_t)cell * 3 + q] = s->geom[(size_t)(s->elem_of_cell[cell] - 1) * 3 + q];
        if ((rc = upload_bulk(h, &S.geom, ge.data(), ge.size()))) return rc;
        // the moduli once per node, if every element at a node holds bitwise the same
        // values there (a medium given as a function of position)
        const int nd = h->ndof;
        std::vector<double> nm((size_t)nd * S.NS, 0.0);
        std::vector<unsigned char> seen(S.NS, 0);
        bool same = true;
        for (int e = 0; e < s->nelem && same; ++e)
          for (int p = 0; p < NP && same; ++p) {
            const size_t g = S.gmap[s->ibool[(size_t)e * NP + p] - 1];
            for (int c = 0; c < nd; ++c) {
              const double x = s->lam_mu[((size_t)e * nd + c) * NP + p];
              double& y = nm[(size_t)c * S.NS + g];
              if (!seen[g]) y = x;
              else if (std::memcmp(&x, &y, sizeof x) != 0) same = false;
            }
            seen[g] = 1;
          }
        if (same && (rc = upload_bulk(h, &S.nodemod, nm.data(), nm.size()))) return rc;
      }
    }
  }
  const size_t nfs = (size_t)h->ndof * h->fstride;
  if ((rc = dev_alloc(h, (char**)&h->d, nfs * rsize(h)))) return rc;
  if ((rc = dev_alloc(h, (char**)&h->v, nfs * rsize(h)))) return rc;
  if ((rc = dev_alloc(h, (char**)&h->a, nfs * rsize(h)))) return rc;
  if ((rc = dev_alloc(h, &h->flags, 4))) return rc;
  if ((rc = dev_alloc(h, &h->info_acc, 2))) return rc;
  if ((rc = dev_alloc(h, &h->info_cnt, 2))) return rc;
  HIPCHK(h, hipMemsetAsync(h->info_acc, 0, 2 * sizeof(unsigned long long), h->stream));
  HIPCHK(h, hipMemsetAsync(h->info_cnt, 0, 2 * sizeof(int), h->stream));
  HIPCHK(h, hipMemsetAsync(h->d, 0, nfs * rsize(h), h->stream));
  HIPCHK(h, hipMemsetAsync(h->v, 0, nfs * rsize(h), h->stream));
  HIPCHK(h, hipMemsetAsync(h->a, 0, nfs * rsize(h), h->stream));
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
  AbsoHost ah;
  ah.node = nd;
  ah.C.assign(C, C + (size_t)h->ndof * n);
  h->abso_host.push_back(ah);
  h->st.ready = false;
  return 0;
}

int sem2d_set_abso_normals(sem2d_solver* h, int n, const double* nrm) {
  if (!h) return 1;
  if (h->abso.empty()) return set_err(h, 3, "set_abso_normals: call sem2d_set_abso first");
  if (h->ndof != 2) return set_err(h, 3, "set_abso_normals: P-SV only");
  if (h->st.on)
    return set_err(h, 3, "set_abso_normals: a boundary that is not flat needs a solver on the general path");
  AbsoDev& ad = h->abso.back();
  if (!nrm || n != ad.n) return set_err(h, 3, "set_abso_normals: size differs from the boundary's");
  return upload(h, &ad.nrm, nrm, (size_t)2 * n);
}

int sem2d_set_abso_stacey(sem2d_solver* h, int nbe, const int* enode, const double* K) {
  if (!h) return 1;
  if (h->abso.empty()) return set_err(h, 3, "set_abso_stacey: call sem2d_set_abso first");
  if (h->ndof != 2) return set_err(h, 3, "set_abso_stacey: P-SV only");
  if (nbe < 1 || !enode || !K) return set_err(h, 3, "set_abso_stacey: empty");
  AbsoHost& ah = h->abso_host.back();
  AbsoDev& ad = h->abso.back();
  if (ad.nbe) return set_err(h, 3, "set_abso_stacey: this boundary already has its Stacey term");
  const int n = (int)ah.node.size();
  const int NG = h->ng;  // edge points per boundary element
  for (int k = 0; k < nbe * NG; ++k)
    if (enode[k] < 0 || enode[k] >= n) return set_err(h, 3, "set_abso_stacey: node out of range");
  ah.senode.assign(enode, enode + (size_t)nbe * NG);
  ah.sK.assign(K, K + (size_t)nbe * NG * 2);
  // general path: global nodes of the edge points, per boundary node its element points
  std::vector<int> eg((size_t)nbe * NG), off(n + 1, 0), idx((size_t)nbe * NG);
  for (int k = 0; k < nbe * NG; ++k) {
    eg[k] = ah.node[enode[k]];
    off[enode[k] + 1]++;
  }
  for (int q = 0; q < n; ++q) off[q + 1] += off[q];
  std::vector<int> fill(off.begin(), off.end() - 1);
  for (int k = 0; k < nbe * NG; ++k) idx[fill[enode[k]]++] = k;  // element order
  int rc;
  if ((rc = upload(h, &ad.enode, eg.data(), eg.size()))) return rc;
  if ((rc = upload(h, &ad.K, K, (size_t)nbe * NG * 2))) return rc;
  if ((rc = upload(h, &ad.soff, off.data(), off.size()))) return rc;
  if ((rc = upload(h, &ad.sidx, idx.data(), idx.size()))) return rc;
  ad.nbe = nbe;
  h->st.ready = false;
  return 0;
}

static int use_general_path(sem2d_solver* h);

int sem2d_set_abso_periodic(sem2d_solver* h) {
  if (!h) return 1;
  if (h->abso.empty()) return set_err(h, 3, "set_abso_periodic: call sem2d_set_abso first");
  AbsoHost& ah = h->abso_host.back();
  AbsoDev& ad = h->abso.back();
  if (!ad.nbe) return set_err(h, 3, "set_abso_periodic: the boundary has no Stacey term");
  if (h->st.eta || h->st.b25)
    return set_err(h, 4, "set_abso_periodic: not with Kelvin-Voigt or 2.5D terms on a structured grid");
  int rc = use_general_path(h);
  if (rc) return rc;
  // both end nodes take the Stacey contributions of the first end node, then the last's
  // (SRC/bc_abso.f90:325-328)
  const int n = (int)ah.node.size(), NG = h->ng, nbe = ad.nbe;
  std::vector<std::vector<int>> lists(n);
  for (int k = 0; k < nbe * NG; ++k) lists[ah.senode[k]].push_back(k);  // element order
  std::vector<int> ends = lists[0];
  ends.insert(ends.end(), lists[n - 1].begin(), lists[n - 1].end());
  lists[0] = ends;
  lists[n - 1] = ends;
  std::vector<int> off(n + 1, 0), idx;
  for (int q = 0; q < n; ++q) {
    idx.insert(idx.end(), lists[q].begin(), lists[q].end());
    off[q + 1] = (int)idx.size();
  }
  if ((rc = upload(h, &ad.soff, off.data(), off.size()))) return rc;
  if ((rc = upload(h, &ad.sidx, idx.data(), idx.size()))) return rc;
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
  h->st.ready = false;
  if (isrc) *isrc = s;
  return 0;
}

int sem2d_set_moment(sem2d_solver* h, int nel, const int* ixi, const int* ieta, const double* cxi,
                     const double* ceta, int* isrc) {
  if (!h) return 1;
  if (nel <= 0 || !ixi || !ieta || !cxi || !ceta) return set_err(h, 3, "set_moment: empty stencil");
  const int s = h->nsrc++;
  const int NG = h->ng;  // stencil points per element line
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
  h->st.ready = false;
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
    if (h->st.on) nd[r] = h->st.gmap[nd[r]];
  }
  int rc = upload(h, &h->rec_node, nd.data(), nd.size());
  if (rc) return rc;
  h->rec_w = nullptr;
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

// the fused tile kernels have neither the viscous nor the 2.5D term: a structured grid
// with either uses the general element path, with the fields in the reference numbering
static int use_general_path(sem2d_solver* h) {
  if (h->st.on) {
    h->st.on = false;
    h->fstride = (size_t)h->npoin;
    const size_t nfs = (size_t)h->ndof * h->fstride * rsize(h);
    HIPCHK(h, hipMemsetAsync(h->d, 0, nfs, h->stream));
    HIPCHK(h, hipMemsetAsync(h->v, 0, nfs, h->stream));
    HIPCHK(h, hipMemsetAsync(h->a, 0, nfs, h->stream));
  }
  return 0;
}

// per-element-point array in the colour-sorted element order of the general path
static int upload_elem_points(sem2d_solver* h, void** dst, const double* src) {
  const int NP = h->np;
  std::vector<double> e((size_t)h->nelem * NP);
  for (int k = 0; k < h->nelem; ++k)
    std::memcpy(&e[(size_t)k * NP], src + (size_t)h->elem_order[k] * NP, NP * sizeof(double));
  return upload_bulk(h, dst, e.data(), e.size());
}

// per-element-point array in cell order for the tile kernels of a structured grid; the
// tile kernels then run pipelined (they hold these terms in that form only)
static int upload_cell_points(sem2d_solver* h, void** dst, const double* src) {
  StructState& S = h->st;
  const int NP = h->np;
  std::vector<double> c((size_t)h->nelem * NP);
  for (int cell = 0; cell < h->nelem; ++cell)
    std::memcpy(&c[(size_t)cell * NP], src + (size_t)S.cell_elem[cell] * NP, NP * sizeof(double));
  int rc = upload_bulk(h, dst, c.data(), c.size());
  if (rc) return rc;
  // one 512-thread workgroup per CU: the variant's LDS and registers allow no more
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  S.pipe_wgs = ((ncu + 7) / 8) * 8;
  S.pipe_threads = 512;
  S.ready = false;
  return 0;
}

int sem2d_set_kelvin_voigt(sem2d_solver* h, const double* eta) {
  if (!h) return 1;
  if (!eta) return set_err(h, 3, "set_kelvin_voigt: null eta");
  if (h->it != 0 || h->nrec || h->flt.on || h->nsrc || !h->abso.empty())
    return set_err(h, 4, "set_kelvin_voigt: call it right after sem2d_create");
  if (h->st.on && (h->nelast == 6 || h->nelast == 2)) return upload_cell_points(h, &h->st.eta, eta);
  int rc = use_general_path(h);
  if (rc) return rc;
  return upload_elem_points(h, &h->kv_eta, eta);
}

int sem2d_set_25d(sem2d_solver* h, const double* beta) {
  if (!h) return 1;
  if (!beta) return set_err(h, 3, "set_25d: null beta");
  if (h->it != 0 || h->nrec || h->flt.on || h->nsrc || !h->abso.empty())
    return set_err(h, 4, "set_25d: call it right after sem2d_create");
  // the general element kernel's copy serves sem2d_compute_fint
  int rc = upload_elem_points(h, &h->beta25, beta);
  if (rc) return rc;
  if (h->st.on && (h->nelast == 6 || h->nelast == 2)) return upload_cell_points(h, &h->st.b25, beta);
  return use_general_path(h);
}

int sem2d_set_periodic(sem2d_solver* h, int n, const int* master, const int* slave) {
  if (!h) return 1;
  if (n <= 0 || !master || !slave) return set_err(h, 3, "set_periodic: bad arguments");
  if (h->it != 0 || h->nrec || h->flt.on || h->nsrc || !h->abso.empty())
    return set_err(h, 4, "set_periodic: call it right after sem2d_create");
  if (h->st.eta || h->st.b25 || h->kv_eta || h->beta25)
    return set_err(h, 4, "set_periodic: call it before sem2d_set_kelvin_voigt and sem2d_set_25d");
  if (h->nper) return set_err(h, 3, "set_periodic: one periodic boundary per solver");
  std::vector<char> mark(h->npoin, 0);
  std::vector<int> m(n), sl(n);
  for (int k = 0; k < n; ++k) {
    m[k] = master[k] - 1;
    sl[k] = slave[k] - 1;
    if (m[k] < 0 || m[k] >= h->npoin || sl[k] < 0 || sl[k] >= h->npoin || m[k] == sl[k] || mark[m[k]] ||
        mark[sl[k]])
      return set_err(h, 3, "set_periodic: nodes out of range or repeated");
    mark[m[k]] = mark[sl[k]] = 1;
  }
  // the pairs act on the reference numbering: the general element path
  int rc = use_general_path(h);
  if (rc) return rc;
  if ((rc = upload(h, &h->per_m, m.data(), n))) return rc;
  if ((rc = upload(h, &h->per_s, sl.data(), n))) return rc;
  h->nper = n;
  return 0;
}

int sem2d_set_receivers_interp(sem2d_solver* h, int nrec, const int* elem, const double* interp, int isamp,
                               int field, int nsamp) {
  if (!h) return 1;
  if (nrec <= 0 || !elem || !interp || isamp < 1 || field < 0 || field > 2 || nsamp < 1)
    return set_err(h, 3, "set_receivers_interp: bad arguments");
  if (h->elem_ibool.empty()) return set_err(h, 4, "set_receivers_interp: element table not kept");
  const int NP = h->np;
  std::vector<int> nd((size_t)nrec * NP);
  for (int r = 0; r < nrec; ++r) {
    if (elem[r] < 1 || elem[r] > h->nelem) return set_err(h, 3, "set_receivers_interp: element out of range");
    for (int k = 0; k < NP; ++k) {
      int g = h->elem_ibool[(size_t)(elem[r] - 1) * NP + k];
      nd[(size_t)r * NP + k] = h->st.on ? h->st.gmap[g] : g;
    }
  }
  int rc = upload(h, &h->rec_node, nd.data(), nd.size());
  if (rc) return rc;
  if ((rc = upload(h, &h->rec_w, interp, (size_t)nrec * NP))) return rc;
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
  if (h->nlsf) return set_err(h, 4, "set_fault: not combined with a linear slip fault");
  if (f->law < 1 || f->law > 3)
    return set_err(h, 3, "set_fault: law must be 1 (SWF), 2 (RSF) or 3 (time weakening, sem2d_set_fault_twf)");
  if ((f->law == 1 && (f->kind < 1 || f->kind > 3)) || (f->law == 2 && (f->kind < 1 || f->kind > 4)))
    return set_err(h, 3, "set_fault: invalid friction kind");
  if (f->oitd < 1 || f->oixd < 1 || f->oix1 < 1 || f->oixn > f->npoin || f->oixn < f->oix1)
    return set_err(h, 3, "set_fault: bad output decimation");
  const int np = f->npoin, nd = h->ndof;
  FaultHost& F = h->flt;
  FaultDev& D = F.dev;
  D.np = np;
  D.ndof = nd;
  // node2 == NULL: a fault with one side, on a domain boundary (symmetry; the
  // reference's bc2 and invM2 are then undefined, SRC/bc_dynflt.f90:256-259)
  const bool two = f->node2 != nullptr;
  if (two && !f->invM2) return set_err(h, 3, "set_fault: invM2 needed with node2");
  std::vector<int> n1(np), n2(np, -1);
  for (int k = 0; k < np; ++k) {
    if (f->node1[k] < 1 || f->node1[k] > h->npoin || (two && (f->node2[k] < 1 || f->node2[k] > h->npoin)))
      return set_err(h, 3, "set_fault: node out of range");
    n1[k] = f->node1[k] - 1;
    if (two) n2[k] = f->node2[k] - 1;
  }
  h->fault_n1 = n1;
  if (two) h->fault_n2 = n2;
  else h->fault_n2.clear();
  h->st.ready = false;
  if (h->st.on)
    for (int k = 0; k < np; ++k) {
      n1[k] = h->st.gmap[n1[k]];
      if (two) n2[k] = h->st.gmap[n2[k]];
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
  std::vector<double> zero_m((size_t)nd * np, 0.0);
  UP(invM2, two ? f->invM2 : zero_m.data(), nd * np);
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
    if (f->law == 3) {  // time weakening alone: set by sem2d_set_fault_twf
      mu[k] = 1.7976931348623157e308;
    } else if (f->law == 1) {
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
  if ((rc = potency_reserve(h, 1024))) return rc;
  record_fault(h);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int sem2d_step(sem2d_solver* h, int nsteps) {
  if (!h) return 1;
  if (nsteps < 0) return set_err(h, 3, "step: negative nsteps");
  return h->fp32 ? do_steps<float>(h, nsteps) : do_steps<double>(h, nsteps);
}

// device field (grid order when structured) -> host array in reference order
static int fetch_field(sem2d_solver* h, double* out, const void* dev) {
  const size_t n = (size_t)h->ndof * h->npoin;
  if (!h->st.on) return read_bulk(h, out, dev, n);
  const size_t S = h->fstride;
  std::vector<double> tmp((size_t)h->ndof * S);
  int rc = read_bulk(h, tmp.data(), dev, tmp.size());
  if (rc) return rc;
  for (int c = 0; c < h->ndof; ++c)
    for (int k = 0; k < h->npoin; ++k) out[(size_t)c * h->npoin + k] = tmp[(size_t)c * S + h->st.gmap[k]];
  return 0;
}

static int store_field(sem2d_solver* h, void* dev, const double* in) {
  const size_t n = (size_t)h->ndof * h->npoin;
  if (!h->st.on) return write_bulk(h, dev, in, n);
  const size_t S = h->fstride;
  std::vector<double> tmp((size_t)h->ndof * S, 0.0);
  for (int c = 0; c < h->ndof; ++c)
    for (int k = 0; k < h->npoin; ++k) tmp[(size_t)c * S + h->st.gmap[k]] = in[(size_t)c * h->npoin + k];
  return write_bulk(h, dev, tmp.data(), tmp.size());
}

int sem2d_fetch_fields(sem2d_solver* h, double* d, double* v, double* a) {
  if (!h) return 1;
  int rc = 0;
  if (d && (rc = fetch_field(h, d, h->d))) return rc;
  if (v && (rc = fetch_field(h, v, h->v))) return rc;
  if (a && (rc = fetch_field(h, a, h->a))) return rc;
  return 0;
}

int sem2d_set_fields(sem2d_solver* h, const double* d, const double* v, const double* a) {
  if (!h) return 1;
  int rc = 0;
  if (d && (rc = store_field(h, h->d, d))) return rc;
  if (v && (rc = store_field(h, h->v, v))) return rc;
  if (a && (rc = store_field(h, h->a, a))) return rc;
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

int sem2d_set_fault_twf(sem2d_solver* h, int kind, const double* par, const double* coord) {
  if (!h) return 1;
  FaultHost& F = h->flt;
  if (!F.on) return set_err(h, 4, "set_fault_twf: call sem2d_set_fault first");
  if (h->it != 0) return set_err(h, 4, "set_fault_twf: the solver has already been stepped");
  if (kind < 1 || kind > 3 || !par || !coord) return set_err(h, 3, "set_fault_twf: bad arguments");
  FaultDev& D = F.dev;
  const int np = D.np;
  D.twf = kind;
  for (int q = 0; q < 9; ++q) D.tw[q] = par[q];
  double* xz = nullptr;
  int rc = upload(h, &xz, coord, (size_t)2 * np);
  if (rc) return rc;
  D.xz = xz;
  // initial friction (SRC/bc_dynflt.f90:404-413): the minimum with the time-weakening
  // coefficient at t = 0, or that coefficient alone (law 3); slip D = 0
  std::vector<double> mu(np);
  HIPCHK(h, hipMemcpy(mu.data(), D.MU, np * sizeof(double), hipMemcpyDeviceToHost));
  for (int k = 0; k < np; ++k) {
    const double mt = twf_mu(kind, par, coord[k], coord[np + k], 0.0, 0.0);
    mu[k] = D.law == 3 ? mt : std::fmin(mu[k], mt);
  }
  HIPCHK(h, hipMemcpy(D.MU, mu.data(), np * sizeof(double), hipMemcpyHostToDevice));
  if (F.taken)  // the friction row of the record of step 0, taken by sem2d_set_fault
    k_fault_record<<<grid_for(F.nout, 64), 64, 0, h->stream>>>(D, F.oix1, F.oixd, F.nout, F.rec);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int sem2d_set_abso_wave(sem2d_solver* h, int n, const double* xz, const double* vdir, const double* tdir,
                        const double* B, const double* par) {
  if (!h) return 1;
  if (h->abso.empty()) return set_err(h, 3, "set_abso_wave: call sem2d_set_abso first");
  AbsoDev& ad = h->abso.back();
  if (n != ad.n || !xz || !vdir || !tdir || !B || !par) return set_err(h, 3, "set_abso_wave: bad arguments");
  if (h->it != 0) return set_err(h, 4, "set_abso_wave: the solver has already been stepped");
  if (h->st.eta || h->st.b25)
    return set_err(h, 4, "set_abso_wave: not with Kelvin-Voigt or 2.5D terms on a structured grid");
  int rc = use_general_path(h);  // the incident field acts on the reference numbering
  if (rc) return rc;
  if ((rc = upload(h, &ad.wxz, xz, (size_t)2 * n))) return rc;
  if ((rc = upload(h, &ad.wV, vdir, (size_t)h->ndof * n))) return rc;
  if ((rc = upload(h, &ad.wT, tdir, (size_t)h->ndof * n))) return rc;
  if ((rc = upload(h, &ad.wB, B, (size_t)n))) return rc;
  ad.wp = WavePar{par[0], par[1], par[2], par[3], par[4], par[5], par[6], par[7]};
  return 0;
}

int sem2d_set_lsf(sem2d_solver* h, int n, const int* node1, const int* node2, const double* nrm, const double* B,
                  const double* W1, const double* W2, double kt, double kn) {
  if (!h) return 1;
  if (n <= 0 || !node1 || !node2 || !nrm || !B || !W1 || !W2) return set_err(h, 3, "set_lsf: bad arguments");
  if (h->nlsf) return set_err(h, 3, "set_lsf: one linear slip fault per solver");
  if (h->it != 0) return set_err(h, 4, "set_lsf: the solver has already been stepped");
  if (h->flt.on || !h->dirs.empty())
    return set_err(h, 4, "set_lsf: not combined with a dynamic fault or Dirichlet boundaries");
  if (h->st.eta || h->st.b25)
    return set_err(h, 4, "set_lsf: not with Kelvin-Voigt or 2.5D terms on a structured grid");
  std::vector<int> a(n), b(n);
  for (int k = 0; k < n; ++k) {
    a[k] = node1[k] - 1;
    b[k] = node2[k] - 1;
    if (a[k] < 0 || a[k] >= h->npoin || b[k] < 0 || b[k] >= h->npoin)
      return set_err(h, 3, "set_lsf: node out of range");
  }
  int rc = use_general_path(h);
  if (rc) return rc;
  const size_t nd = (size_t)h->ndof * n;
  if ((rc = upload(h, &h->lsf_n1, a.data(), n)) || (rc = upload(h, &h->lsf_n2, b.data(), n)) ||
      (rc = upload(h, &h->lsf_nrm, nrm, (size_t)2 * n)) || (rc = upload(h, &h->lsf_B, B, nd)) ||
      (rc = upload(h, &h->lsf_W1, W1, nd)) || (rc = upload(h, &h->lsf_W2, W2, nd)))
    return rc;
  h->lsf_kt = kt;
  h->lsf_kn = kn;
  h->nlsf = n;
  return 0;
}

int sem2d_set_hht(sem2d_solver* h, double alpha) {
  if (!h) return 1;
  if (!(alpha >= 0.0 && alpha <= 1.0)) return set_err(h, 3, "set_hht: alpha must lie in [0, 1]");
  if (h->it != 0 || h->nrec || h->flt.on || h->nsrc || !h->abso.empty())
    return set_err(h, 4, "set_hht: call it right after sem2d_create");
  if (h->st.eta || h->st.b25 || h->kv_eta || h->beta25)
    return set_err(h, 4, "set_hht: call it before sem2d_set_kelvin_voigt and sem2d_set_25d");
  if (h->leapfrog) return set_err(h, 4, "set_hht: the solver is set to leapfrog");
  // the fields at n + alpha live in the reference numbering: the general element path
  int rc = use_general_path(h);
  if (rc) return rc;
  if (!h->da) {
    const size_t n = (size_t)h->ndof * h->npoin;
    if (h->fp32) {
      float *p1, *p2;
      if ((rc = dev_alloc(h, &p1, n)) || (rc = dev_alloc(h, &p2, n))) return rc;
      h->da = p1;
      h->va = p2;
    } else {
      double *p1, *p2;
      if ((rc = dev_alloc(h, &p1, n)) || (rc = dev_alloc(h, &p2, n))) return rc;
      h->da = p1;
      h->va = p2;
    }
  }
  h->hht = 1;
  h->alpha = alpha;
  return 0;
}

int sem2d_set_fault_normal(sem2d_solver* h, int kind, double coef, double v) {
  if (!h) return 1;
  FaultHost& F = h->flt;
  if (!F.on) return set_err(h, 4, "set_fault_normal: call sem2d_set_fault first");
  if (h->it != 0) return set_err(h, 4, "set_fault_normal: the solver has already been stepped");
  if (kind < 0 || kind > 3) return set_err(h, 3, "set_fault_normal: kind must be 0 to 3");
  F.dev.nor_kind = kind;
  F.dev.nor_coef = coef;
  F.dev.nor_v = v;
  return 0;
}

int sem2d_set_fault_sides(sem2d_solver* h, int on) {
  if (!h) return 1;
  FaultHost& F = h->flt;
  if (!F.on) return set_err(h, 4, "set_fault_sides: call sem2d_set_fault first");
  if (h->it != 0) return set_err(h, 4, "set_fault_sides: the solver has already been stepped");
  if (!on) {
    F.srec = nullptr;  // the block stays with the handle's allocations
    return 0;
  }
  int rc = dev_alloc(h, &F.srec, (size_t)F.nsamp * 4 * h->ndof * F.nout);
  if (rc) return rc;
  if (F.taken) record_sides(h, 0);  // the record of step 0, taken by sem2d_set_fault
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int sem2d_fetch_fault_sides(sem2d_solver* h, double* out, int* nsamp_taken) {
  if (!h) return 1;
  FaultHost& F = h->flt;
  if (!F.on || !F.srec) return set_err(h, 4, "fetch_fault_sides: no per-side records (sem2d_set_fault_sides)");
  if (out)
    HIPCHK(h, hipMemcpy(out, F.srec, (size_t)F.taken * 4 * h->ndof * F.nout * sizeof(double),
                        hipMemcpyDeviceToHost));
  if (nsamp_taken) *nsamp_taken = F.taken;
  return 0;
}

int sem2d_fetch_potency(sem2d_solver* h, double* out, int* nrows) {
  if (!h) return 1;
  if (!h->flt.on) return set_err(h, 4, "fetch_potency: no fault");
  if (out && h->flt.pot_rows)
    HIPCHK(h, hipMemcpy(out, h->flt.pot, (size_t)h->flt.pot_rows * 2 * (h->ndof + 1) * sizeof(double),
                        hipMemcpyDeviceToHost));
  if (nrows) *nrows = h->flt.pot_rows;
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

int sem2d_compute_fint_noalloc(sem2d_solver* h, const double* d, double* f) {
  if (!h) return 1;
  if (!d || !f) return set_err(h, 3, "compute_fint_noalloc: null array");
  const size_t n = (size_t)h->ndof * h->npoin;
  // 16 bytes of scratch per value.  FP64 fields: d at [0, 8n), f at [8n, 16n).  FP32
  // fields: the host d as doubles at [0, 8n), d as floats at [8n, 12n), f as floats at
  // [12n, 16n), then f as doubles back at [0, 8n) (the double d is consumed by then).
  if (n > NOALLOC_SCRATCH_BYTES / 16)
    return set_err(h, 3, "compute_fint_noalloc: ndof*npoin = %zu exceeds the scratch limit of %zu values",
                   n, NOALLOC_SCRATCH_BYTES / 16);
  unsigned char* base = (unsigned char*)scratch_on_device(h);
  if (!base) return 10;
  double* dd = (double*)base;
  const int nb = grid_for(n, 256);
  HIPCHK(h, hipMemcpyAsync(dd, d, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
  double* out;
  if (h->fp32) {
    float* dr = (float*)(base + 8 * n);
    float* fr = (float*)(base + 12 * n);
    k_to_float<<<nb, 256, 0, h->stream>>>(dd, dr, n);
    k_fill<float><<<nb, 256, 0, h->stream>>>(fr, 0.0f, n);
    launch_force<float>(h, dr, fr);
    k_to_double<<<nb, 256, 0, h->stream>>>(fr, dd, n);
    out = dd;
  } else {
    double* fd = (double*)(base + 8 * n);
    k_fill<double><<<nb, 256, 0, h->stream>>>(fd, 0.0, n);
    launch_force<double>(h, dd, fd);
    out = fd;
  }
  if (hipGetLastError() != hipSuccess) return set_err(h, 10, "compute_fint_noalloc: kernel launch failed");
  HIPCHK(h, hipMemcpyAsync(f, out, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int sem2d_set_dirichlet(sem2d_solver* h, int n, const int* node, int comp) {
  if (!h) return 1;
  if (n <= 0 || !node) return set_err(h, 3, "set_dirichlet: empty boundary");
  if (comp < 1 || comp > 3 || (h->ndof == 1 && comp != 1))
    return set_err(h, 3, "set_dirichlet: comp must be 1 (component 1), 2 (component 2) or 3 (both; P-SV)");
  if (h->nlsf) return set_err(h, 4, "set_dirichlet: not combined with a linear slip fault");
  sem2d_solver::DirBC D;
  D.node.resize(n);
  for (int k = 0; k < n; ++k) {
    if (node[k] < 1 || node[k] > h->npoin) return set_err(h, 3, "set_dirichlet: node out of range");
    D.node[k] = node[k] - 1;
  }
  D.comp = comp;
  D.after_fault = h->flt.on;
  std::vector<int> dn(D.node);
  if (h->st.on)
    for (int& k : dn) k = h->st.gmap[k];
  int rc = upload(h, &D.node_dev, dn.data(), dn.size());
  if (rc) return rc;
  h->dirs.push_back(std::move(D));
  h->st.ready = false;
  return 0;
}

int sem2d_set_time_scheme(sem2d_solver* h, int kind) {
  if (!h) return 1;
  if (kind != 0 && kind != 1) return set_err(h, 3, "set_time_scheme: kind must be 0 (newmark) or 1 (leapfrog)");
  if (kind == 1 && (h->beta != 0.0 || h->gamma != 0.5))
    return set_err(h, 3, "set_time_scheme: leapfrog needs beta = 0, gamma = 1/2 in the descriptor");
  if (h->it != 0) return set_err(h, 4, "set_time_scheme: the solver has already been stepped");
  if (kind == 1 && h->hht) return set_err(h, 4, "set_time_scheme: the solver is set to HHT-alpha");
  h->leapfrog = kind;
  return 0;
}

int sem2d_set_itinfo(sem2d_solver* h, int itinfo) {
  if (!h) return 1;
  if (itinfo < 0) return set_err(h, 3, "set_itinfo: negative interval");
  h->itinfo = itinfo;
  return 0;
}

int sem2d_field_info(sem2d_solver* h, double* vmax, double* dmax, int* nonfinite) {
  if (!h) return 1;
  const size_t n = (size_t)h->ndof * h->fstride;
  const int nb = (int)std::min<size_t>((n + 255) / 256, 4096);
  if (h->fp32)
    k_field_max<float><<<nb, 256, 0, h->stream>>>(n, (const float*)h->d, (const float*)h->v, (const float*)h->a,
                                                 h->info_acc, h->info_cnt);
  else
    k_field_max<double><<<nb, 256, 0, h->stream>>>(n, (const double*)h->d, (const double*)h->v,
                                                  (const double*)h->a, h->info_acc, h->info_cnt);
  unsigned long long acc[2];
  int cnt[2];
  HIPCHK(h, hipMemcpyAsync(acc, h->info_acc, sizeof acc, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(cnt, h->info_cnt, sizeof cnt, hipMemcpyDeviceToHost, h->stream));
  k_info_row<<<1, 64, 0, h->stream>>>(h->it, h->it * h->dt, h->info_acc, h->info_cnt, nullptr);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  double x[2];
  std::memcpy(x, acc, sizeof x);
  if (vmax) *vmax = x[0];
  if (dmax) *dmax = x[1];
  if (nonfinite) *nonfinite = cnt[0];
  return 0;
}

int sem2d_fetch_info(sem2d_solver* h, double* rows, int* nrows) {
  if (!h) return 1;
  if (rows && h->info_rows)
    HIPCHK(h, hipMemcpy(rows, h->info, (size_t)h->info_rows * 4 * sizeof(double), hipMemcpyDeviceToHost));
  if (nrows) *nrows = h->info_rows;
  return 0;
}

int sem2d_set_tile_config(sem2d_solver* h, int threads, int wg_per_cu) {
  if (!h) return 1;
  if (!h->st.on) return set_err(h, 4, "set_tile_config: not a structured (MESH_CART) grid");
  if ((threads != 256 && threads != 512) || wg_per_cu > 4)
    return set_err(h, 3, "set_tile_config: threads must be 256 or 512 and wg_per_cu <= 4");
  StructState& S = h->st;
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  S.pipe_threads = threads;
  if (wg_per_cu < 0) {  // automatic: the measured defaults
    const int pc = h->fp32 ? 3 : 2;
    S.pipe_wgs = S.ntiles >= 4 * ncu * pc ? ncu * pc : 0;
  }
  else S.pipe_wgs = ncu * wg_per_cu;
  S.pipe_wgs = ((S.pipe_wgs + 7) / 8) * 8;
  return 0;
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
  if (h->adj) adj_free(h->adj);
  if (h->flt.pot) hipFree(h->flt.pot);
  if (h->info) hipFree(h->info);
  for (void* p : h->allocs) hipFree(p);
  for (void* p : h->st.prep_allocs) hipFree(p);
  for (auto& b : h->abso) { (void)b; }
  if (h->src_dev) hipFree(h->src_dev);
  if (h->ampl_dev) hipFree(h->ampl_dev);
  if (h->ev0) hipEventDestroy(h->ev0);
  if (h->ev1) hipEventDestroy(h->ev1);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
}

}  // extern "C"

// ===========================================================================
// Batched ensemble: many independent structured scenarios advanced together.
//
// The member solvers' grids are laid side by side in one set of field arrays
// (member m's storage block starts at node_base[m]; component stride = the sum of
// the members' strides).  Tiles, deferred nodes, sources, absorbing coefficients
// and fault pairs of all members are concatenated with their indices offset, so
// one persistent tile launch and one deferred launch cover every scenario.  The
// step index lives in device memory (it_dev), which lets a captured hipGraph of a
// chunk of steps be replayed for the whole time loop.  Every member performs
// exactly the arithmetic of its own solver, so batch and one-by-one results are
// bitwise identical.
// ===========================================================================
namespace {

template <typename R>
__global__ void k_batch_record(const int* __restrict__ itp, int nrec, int ndof, const int* __restrict__ rec_node,
                               const R* __restrict__ field, size_t fstride, int isamp, int nsamp,
                               double* __restrict__ seis, FaultDev F, const int* __restrict__ fidx, int nout,
                               int oit0, int oitd, int fnsamp, double* __restrict__ frec) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int it = *itp;
  if (q < nrec && it % isamp == 0 && it / isamp < nsamp) {
    double* out = seis + (size_t)(it / isamp) * ndof * nrec;
    for (int c = 0; c < ndof; ++c) out[(size_t)c * nrec + q] = (double)field[(size_t)c * fstride + rec_node[q]];
  }
  if (q < nout && it >= oit0 && (it - oit0) % oitd == 0 && (it - oit0) / oitd < fnsamp) {
    double* out = frec + (size_t)((it - oit0) / oitd) * 6 * nout;
    const int k = fidx[q];
    out[0 * nout + q] = F.D[k];
    out[1 * nout + q] = F.V[k];
    out[2 * nout + q] = F.T[k];
    out[3 * nout + q] = F.T[F.np + k];
    out[4 * nout + q] = F.MU[k];
    out[5 * nout + q] = F.Tstick[k];
  }
}

__global__ void k_tick(int* itp) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *itp = *itp + 1;
}

}  // namespace

struct sem2d_batch {
  int n = 0;
  std::vector<sem2d_solver*> mem;
  int ndof = 0, nelast = 0, coef_mode = 0, fp32 = 0;
  double dt = 0, beta = 0, gamma = 0;
  int leapfrog = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  size_t fstride = 0;
  std::vector<size_t> node_base;
  std::vector<int> fault_base, rec_base, fout_base, src_base;
  StepDev X;
  int ncu = 256;
  // records
  int nrec = 0, isamp = 1, field = 1, nsamp = 0;
  int* rec_node = nullptr;
  double* seis = nullptr;
  int fnout = 0, oit0 = 0, oitd = 1, fnsamp = 0;
  int* fidx = nullptr;
  double* frec = nullptr;
  int* it_dev = nullptr;
  int* flags = nullptr;
  int it = 0;
  // potency rows: pot[member][pot_cap][2*(ndof+1)]; the row of the state after step s
  // is pot_shift + s (pot_shift + 1 rows precede the batch's first step)
  int* fbase_dev = nullptr;  // [n+1] first fault pair of each member
  double* pot = nullptr;
  int pot_cap = 0, pot_shift = 0;
  // graph of a chunk of steps
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int gsteps = 0;
  double last_ms = 0.0;
  std::string err;
  std::vector<void*> allocs;
};

namespace {

int berr(sem2d_batch* b, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (b) b->err = buf;
  return code;
}

#define BHIPCHK(b, call)                                                                  \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return berr(b, 10, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

template <typename T>
int balloc(sem2d_batch* b, T** p, size_t n) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) return berr(b, 11, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
  b->allocs.push_back(q);
  *p = (T*)q;
  return 0;
}

template <typename T>
int bupload(sem2d_batch* b, T** p, const std::vector<T>& src) {
  int rc = balloc(b, p, src.size());
  if (rc) return rc;
  if (!src.empty())
    BHIPCHK(b, hipMemcpyAsync(*p, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, b->stream));
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  return 0;
}

int bupload_bulk(sem2d_batch* b, void** p, const std::vector<double>& src) {
  if (!b->fp32) return bupload(b, (double**)p, src);
  std::vector<float> tmp(src.begin(), src.end());
  return bupload(b, (float**)p, tmp);
}

// device -> host copy of n doubles
int bdown(sem2d_batch* b, std::vector<double>& dst, const double* src, size_t n) {
  dst.resize(n);
  if (n) BHIPCHK(b, hipMemcpy(dst.data(), src, n * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int check_members(sem2d_batch* b) {
  sem2d_solver* h0 = b->mem[0];
  for (int m = 0; m < b->n; ++m) {
    sem2d_solver* h = b->mem[m];
    if (!h) return berr(b, 3, "member %d is null", m);
    if (!h->st.on) return berr(b, 3, "member %d is not a structured (MESH_CART) grid", m);
    if (h->it != 0) return berr(b, 3, "member %d has already been stepped", m);
    if (h->ndof != h0->ndof || h->nelast != h0->nelast || h->coef_mode != h0->coef_mode || h->fp32 != h0->fp32)
      return berr(b, 3, "member %d: ndof, element kernel, coefficient mode and bulk precision must match", m);
    if (h->dt != h0->dt || h->beta != h0->beta || h->gamma != h0->gamma || h->leapfrog != h0->leapfrog)
      return berr(b, 3, "member %d: time step, time scheme and Newmark parameters must match", m);
    if (h->flt.on != h0->flt.on) return berr(b, 3, "member %d: either every member has a fault or none", m);
    if (h->flt.srec) return berr(b, 3, "member %d: per-side fault records (osides) are not batched", m);
    if (h->flt.on && h->flt.dev.twf) return berr(b, 3, "member %d: time-weakening friction is not batched", m);
    if (h->flt.on && h->flt.dev.nor_kind != 1)
      return berr(b, 3, "member %d: normal stress laws other than Coulomb are not batched", m);
    for (const AbsoHost& B : h->abso_host)
      if (!B.senode.empty()) return berr(b, 3, "member %d: Stacey boundaries are not batched", m);
    if (h->st.eta || h->st.b25) return berr(b, 3, "member %d: Kelvin-Voigt and 2.5D terms are not batched", m);
    if (h->flt.on && h->fault_n2.empty())
      return berr(b, 3, "member %d: faults with one side (on a domain boundary) are not batched", m);
    if (!h->dirs.empty()) return berr(b, 3, "member %d: Dirichlet boundaries are not batched", m);
    if (h->flt.on) {
      const FaultDev &F = h->flt.dev, &F0 = h0->flt.dev;
      if (F.law != F0.law || F.kind != F0.kind || F.opening != F0.opening || F.healing != F0.healing ||
          F.a2v != F0.a2v || F.a2d != F0.a2d)
        return berr(b, 3, "member %d: fault law, kind, opening, healing and time coefficients must match", m);
      const int oit0 = h->flt.oit - h->flt.taken * h->flt.oitd;
      const int oit00 = h0->flt.oit - h0->flt.taken * h0->flt.oitd;
      if (h->flt.oitd != h0->flt.oitd || oit0 != oit00)
        return berr(b, 3, "member %d: fault output times must match", m);
    }
  }
  return 0;
}

void batch_drop_graph(sem2d_batch* b);

// room for `rows` potency rows per member (outside any captured graph: a new buffer
// drops the graph, whose launches hold the old address).  The rows already recorded
// (pot_shift + 1 + it) move over; the first reservation takes them from the members.
int batch_pot_reserve(sem2d_batch* b, int rows) {
  if (rows <= b->pot_cap) return 0;
  const int cap = std::max(rows, 2 * b->pot_cap);
  const size_t nc = 2 * (size_t)(b->ndof + 1);
  const int have = b->pot ? b->pot_shift + 1 + b->it : b->pot_shift + 1;
  double* p = nullptr;
  BHIPCHK(b, hipMalloc(&p, (size_t)b->n * cap * nc * sizeof(double)));
  for (int m = 0; m < b->n; ++m) {
    const double* src = b->pot ? b->pot + (size_t)m * b->pot_cap * nc : b->mem[m]->flt.pot;
    if (have > 0)
      BHIPCHK(b, hipMemcpyAsync(p + (size_t)m * cap * nc, src, (size_t)have * nc * sizeof(double),
                                hipMemcpyDeviceToDevice, b->stream));
  }
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  if (b->pot) hipFree(b->pot);
  b->pot = p;
  b->pot_cap = cap;
  batch_drop_graph(b);
  return 0;
}

// concatenate the device-resident per-member arrays that only exist on the device
int gather_fault(sem2d_batch* b) {
  const int nd = b->ndof;
  int np = 0;
  b->fault_base.assign(b->n, 0);
  for (int m = 0; m < b->n; ++m) {
    b->fault_base[m] = np;
    np += b->mem[m]->flt.dev.np;
  }
  FaultDev F = b->mem[0]->flt.dev;
  F.np = np;
  std::vector<int> n1, n2;
  std::vector<double> cat[22];
  // arrays and their row counts ([rows][np] layout)
  struct Arr { const double* FaultDev::*p; int rows; };
  const Arr arrs[] = {{&FaultDev::nrm, 2}, {&FaultDev::B, nd},    {&FaultDev::invM1, nd}, {&FaultDev::invM2, nd},
                      {&FaultDev::Z, nd},  {&FaultDev::T0, 2},    {&FaultDev::coh, 1},    {&FaultDev::p0, 1},
                      {&FaultDev::p1, 1},  {&FaultDev::p2, 1},    {&FaultDev::p3, 1},     {&FaultDev::p4, 1},
                      {&FaultDev::p5, 1},  {&FaultDev::Tc, 1},    {&FaultDev::coeft, 1}};
  struct State { double* FaultDev::*p; int rows; };
  const State states[] = {{&FaultDev::theta, 1}, {&FaultDev::D, nd},     {&FaultDev::V, nd}, {&FaultDev::T, 2},
                          {&FaultDev::Tstick, 2}, {&FaultDev::MU, 1},    {&FaultDev::sigma, 1}};
  const int na = sizeof(arrs) / sizeof(arrs[0]), ns = sizeof(states) / sizeof(states[0]);
  for (int q = 0; q < na; ++q) cat[q].assign((size_t)arrs[q].rows * np, 0.0);
  for (int q = 0; q < ns; ++q) cat[na + q].assign((size_t)states[q].rows * np, 0.0);
  for (int m = 0; m < b->n; ++m) {
    const FaultDev& Fm = b->mem[m]->flt.dev;
    const int nm = Fm.np, o = b->fault_base[m];
    std::vector<int> t1(nm), t2(nm);
    BHIPCHK(b, hipMemcpy(t1.data(), Fm.n1, nm * sizeof(int), hipMemcpyDeviceToHost));
    BHIPCHK(b, hipMemcpy(t2.data(), Fm.n2, nm * sizeof(int), hipMemcpyDeviceToHost));
    for (int k = 0; k < nm; ++k) {
      n1.push_back((int)(t1[k] + b->node_base[m]));
      n2.push_back((int)(t2[k] + b->node_base[m]));
    }
    std::vector<double> tmp;
    for (int q = 0; q < na + ns; ++q) {
      const int rows = q < na ? arrs[q].rows : states[q - na].rows;
      const double* src = q < na ? Fm.*(arrs[q].p) : Fm.*(states[q - na].p);
      int rc = bdown(b, tmp, src, (size_t)rows * nm);
      if (rc) return rc;
      for (int r = 0; r < rows; ++r)
        std::copy(tmp.begin() + (size_t)r * nm, tmp.begin() + (size_t)(r + 1) * nm, cat[q].begin() + (size_t)r * np + o);
    }
  }
  int rc;
  int* ip;
  if ((rc = bupload(b, &ip, n1))) return rc;
  F.n1 = ip;
  if ((rc = bupload(b, &ip, n2))) return rc;
  F.n2 = ip;
  for (int q = 0; q < na + ns; ++q) {
    double* dp;
    if ((rc = bupload(b, &dp, cat[q]))) return rc;
    if (q < na) F.*(arrs[q].p) = dp;
    else F.*(states[q - na].p) = dp;
  }
  F.nr_fail = b->flags;
  b->X.fault_on = true;
  b->X.F = F;
  // potency: member pair ranges, and the rows recorded before the batch (one per member
  // at fault set-up; equal for all members, none of which has been stepped)
  std::vector<int> fb(b->fault_base);
  fb.push_back(np);
  if ((rc = bupload(b, &b->fbase_dev, fb))) return rc;
  const int r0 = b->mem[0]->flt.pot_rows;
  for (int m = 0; m < b->n; ++m)
    if (b->mem[m]->flt.pot_rows != r0) return berr(b, 3, "member %d: potency rows differ from member 0's", m);
  if (r0 < 1) return berr(b, 3, "batch: members have no potency row of step 0");
  b->pot_shift = r0 - 1;
  if ((rc = batch_pot_reserve(b, r0))) return rc;
  // fault records: output nodes oix1:oixn:oixd of every member, member after member
  std::vector<int> fidx;
  b->fout_base.assign(b->n, 0);
  b->fnsamp = 0;
  for (int m = 0; m < b->n; ++m) {
    const FaultHost& H = b->mem[m]->flt;
    b->fout_base[m] = (int)fidx.size();
    for (int o = 0; o < H.nout; ++o) fidx.push_back(b->fault_base[m] + H.oix1 - 1 + o * H.oixd);
    b->fnsamp = std::max(b->fnsamp, H.nsamp);
  }
  b->fnout = (int)fidx.size();
  const FaultHost& H0 = b->mem[0]->flt;
  b->oitd = H0.oitd;
  b->oit0 = H0.oit - H0.taken * H0.oitd;
  if ((rc = bupload(b, &b->fidx, fidx))) return rc;
  if ((rc = balloc(b, &b->frec, (size_t)b->fnsamp * 6 * std::max(b->fnout, 1)))) return rc;
  return 0;
}

template <typename R>
void launch_batch_record(sem2d_batch* b) {
  const int nq = std::max(std::max(b->nrec, b->fnout), 1);
  const R* fld = (const R*)(b->field == 0 ? b->X.d : (b->field == 1 ? b->X.v : b->X.a));
  k_batch_record<R><<<grid_for(nq, 64), 64, 0, b->stream>>>(b->it_dev, b->nrec, b->ndof, b->rec_node, fld,
                                                            b->fstride, b->isamp, b->nsamp, b->seis, b->X.F,
                                                            b->fidx, b->fnout, b->oit0, b->oitd, b->fnsamp,
                                                            b->frec);
}

template <typename R>
void launch_batch_steps(sem2d_batch* b, int nsteps) {
  const StepCoefs<R> K = step_coefs<R>(b->dt, b->beta, b->gamma, b->leapfrog);
  const R dt = K.dt, c1 = K.c1, c2 = K.c2, c3 = K.c3, c4 = K.c4;
  for (int s = 0; s < nsteps; ++s) {
    launch_structured_step<R>(b->X, 0, dt, c1, c2, c3, c4);
    if (b->X.fault_on)
      k_batch_potency<R><<<b->n, 256, 0, b->stream>>>(b->X.F, b->fbase_dev, (const R*)b->X.d, (const R*)b->X.v,
                                                      b->fstride, b->it_dev, b->pot_shift, b->pot_cap, b->pot);
    launch_batch_record<R>(b);
    k_tick<<<1, 64, 0, b->stream>>>(b->it_dev);
  }
}

void batch_drop_graph(sem2d_batch* b) {
  if (b->gexec) hipGraphExecDestroy(b->gexec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
  b->gsteps = 0;
}

int batch_graph(sem2d_batch* b, int chunk) {
  if (b->gexec && b->gsteps == chunk) return 0;
  batch_drop_graph(b);
  BHIPCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  if (b->fp32) launch_batch_steps<float>(b, chunk);
  else launch_batch_steps<double>(b, chunk);
  BHIPCHK(b, hipStreamEndCapture(b->stream, &b->graph));
  BHIPCHK(b, hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
  b->gsteps = chunk;
  return 0;
}

// launch configuration of the tile kernel: threads per workgroup (256: 10-element
// batches, 512: 20-element batches) and workgroups per CU (0: one workgroup per tile)
void batch_set_config(sem2d_batch* b, int threads, int per_cu) {
  b->X.pipe_threads = threads;
  b->X.pipe_wgs = per_cu > 0 ? ((b->ncu * per_cu + 7) / 8) * 8 : 0;
  batch_drop_graph(b);
}

int batch_run(sem2d_batch* b, int nsteps, bool use_graph) {
  const int chunk = std::min(nsteps, 64);
  if (b->X.fault_on) {
    int rc = batch_pot_reserve(b, b->pot_shift + 1 + b->it + nsteps);
    if (rc) return rc;
  }
  BHIPCHK(b, hipEventRecord(b->ev0, b->stream));
  int done = 0;
  if (use_graph && chunk > 0) {
    int rc = batch_graph(b, chunk);
    if (rc) return rc;
    for (; done + chunk <= nsteps; done += chunk) BHIPCHK(b, hipGraphLaunch(b->gexec, b->stream));
  }
  if (done < nsteps) {
    if (b->fp32) launch_batch_steps<float>(b, nsteps - done);
    else launch_batch_steps<double>(b, nsteps - done);
  }
  BHIPCHK(b, hipEventRecord(b->ev1, b->stream));
  BHIPCHK(b, hipGetLastError());
  BHIPCHK(b, hipEventSynchronize(b->ev1));
  float ms = 0.f;
  BHIPCHK(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = nsteps > 0 ? ms / nsteps : 0.0;
  b->it += nsteps;
  if (b->X.fault_on) {
    int flag = 0;
    BHIPCHK(b, hipMemcpy(&flag, b->flags, sizeof(int), hipMemcpyDeviceToHost));
    if (flag) return berr(b, 20, "NR_Solver has exceeded the maximum iterations (200)");
  }
  return 0;
}

int rec_taken(const sem2d_batch* b) { return b->nrec ? std::min(b->nsamp, b->it / b->isamp + 1) : 0; }
int flt_taken(const sem2d_batch* b) {
  if (!b->X.fault_on || b->it < b->oit0) return 0;
  return std::min(b->fnsamp, (b->it - b->oit0) / b->oitd + 1);
}

}  // namespace

extern "C" {

int sem2d_batch_create(int n, sem2d_solver* const* members, sem2d_batch** out) {
  if (!out) return 1;
  *out = nullptr;
  if (n <= 0 || !members) return 2;
  sem2d_batch* b = new sem2d_batch();
  *out = b;
  b->n = n;
  b->mem.assign(members, members + n);
  int rc = check_members(b);
  if (rc) return rc;
  sem2d_solver* h0 = b->mem[0];
  b->ndof = h0->ndof;
  b->nelast = h0->nelast;
  b->coef_mode = h0->coef_mode;
  b->fp32 = h0->fp32;
  b->dt = h0->dt;
  b->beta = h0->beta;
  b->gamma = h0->gamma;
  b->leapfrog = h0->leapfrog;
  const int nd = b->ndof;
  const size_t rs = b->fp32 ? sizeof(float) : sizeof(double);
  BHIPCHK(b, hipStreamCreate(&b->stream));
  BHIPCHK(b, hipEventCreate(&b->ev0));
  BHIPCHK(b, hipEventCreate(&b->ev1));
  {
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, dev);
  }
  // member tables, concatenated
  StructTables A;
  std::vector<double> rmdef_c[2], frm_c[2], frm2_c[2];  // fault: inverse mass of node 1 / 2
  std::vector<SrcEntry> dsrc;
  std::vector<double> ampl;
  int ampl_nt = 0, nsrc = 0;
  for (int m = 0; m < n; ++m) {
    sem2d_solver* h = b->mem[m];
    if ((rc = sync_sources(h))) return berr(b, rc, "member %d: %s", m, h->err.c_str());
    ampl_nt = std::max(ampl_nt, h->ampl_nt);
  }
  ampl_nt = std::max(ampl_nt, 1);  // every source owns a row of the table
  b->node_base.assign(n, 0);
  b->src_base.assign(n, 0);
  size_t nb = 0, cb = 0, shared_off = 0;
  for (int m = 0; m < n; ++m) {
    sem2d_solver* h = b->mem[m];
    StructTables X;
    if ((rc = struct_tables(h, X))) return berr(b, rc, "member %d: %s", m, h->err.c_str());
    b->node_base[m] = nb;
    const int tb = (int)A.tiles.size(), pb = (int)A.part_off.size(), sb = (int)A.dsrc.size();
    const int ab = (int)(A.dabso.size() / nd);
    const int cellw = b->coef_mode == COEF_SHARED ? (int)shared_off : (int)cb;
    if (nb + h->st.NS > 0x7fffffffULL || cb + h->nelem > 0x7fffffffULL) return berr(b, 5, "batch too large");
    for (size_t t = 0; t < X.tiles.size(); ++t) {
      A.tiles.push_back(X.tiles[t]);
      A.grids.push_back(make_int4((int)nb, X.grids[t].y, X.grids[t].z, cellw));
    }
    A.spec_cnt.insert(A.spec_cnt.end(), X.spec_cnt.begin(), X.spec_cnt.end());
    A.spec_idx.insert(A.spec_idx.end(), X.spec_idx.begin(), X.spec_idx.end());
    for (DefNode d : X.defs) {
      d.g += (int)nb;
      d.pstart += pb;
      d.sstart += sb;
      d.astart += ab;
      A.defs.push_back(d);
    }
    for (int po : X.part_off) A.part_off.push_back(po + tb * SLOTS);
    for (SrcEntry e : X.dsrc) {
      e.src += nsrc;
      A.dsrc.push_back(e);
    }
    A.dabso.insert(A.dabso.end(), X.dabso.begin(), X.dabso.end());
    const size_t ndef = X.defs.size(), nf = X.fault_g.size();
    for (int c = 0; c < nd; ++c) {
      rmdef_c[c].insert(rmdef_c[c].end(), X.rmdef.begin() + c * ndef, X.rmdef.begin() + (c + 1) * ndef);
    }
    (void)nf;
    if (h->flt.on) {  // per-pair inverse masses, in the batch's pair order
      std::vector<double> r1, r2;
      split_fault_rm(X, nd, h->fault_n1.size(), r1, r2);
      const size_t npm = r1.size() / nd;
      for (int c = 0; c < nd; ++c) {
        frm_c[c].insert(frm_c[c].end(), r1.begin() + c * npm, r1.begin() + (c + 1) * npm);
        frm2_c[c].insert(frm2_c[c].end(), r2.begin() + c * npm, r2.begin() + (c + 1) * npm);
      }
    }
    // source amplitude rows, padded to the common length
    b->src_base[m] = nsrc;
    for (int s = 0; s < h->nsrc; ++s) {
      std::vector<double> row(ampl_nt, 0.0);
      std::copy(h->stf[s].begin(), h->stf[s].end(), row.begin());
      ampl.insert(ampl.end(), row.begin(), row.end());
    }
    nsrc += h->nsrc;
    nb += h->st.NS;
    cb += h->nelem;
    shared_off += (size_t)h->nelast * NP;
  }
  b->fstride = nb;
  for (int c = 0; c < nd; ++c) {
    A.rmdef.insert(A.rmdef.end(), rmdef_c[c].begin(), rmdef_c[c].end());
  }
  StepDev& X = b->X;
  X.ndof = nd;
  X.nelast = b->nelast;
  X.coef_mode = b->coef_mode;
  X.fstride = nb;
  X.stream = b->stream;
  const int ntiles = (int)A.tiles.size();
  int4 *tiles, *grids;
  int *sc, *si, *po;
  DefDev* defs;
  int2* defx;
  SrcEntry* ds;
  double *dab, *amp;
  if ((rc = bupload(b, &tiles, A.tiles))) return rc;
  if ((rc = bupload(b, &grids, A.grids))) return rc;
  if ((rc = bupload(b, &sc, A.spec_cnt))) return rc;
  if ((rc = bupload(b, &si, A.spec_idx))) return rc;
  {
    std::vector<DefDev> dn;
    std::vector<int2> dx;
    if (!pack_defs(A.defs, A.part_off, dn, dx)) return berr(b, 5, "batch: more than 255 entries at a deferred node");
    if ((rc = bupload(b, &defs, dn))) return rc;
    if ((rc = bupload(b, &defx, dx))) return rc;
  }
  if ((rc = bupload(b, &po, A.part_off))) return rc;
  if ((rc = bupload(b, &ds, A.dsrc))) return rc;
  if ((rc = bupload(b, &dab, A.dabso))) return rc;
  if (ampl.empty()) ampl.assign(1, 0.0);
  if ((rc = bupload(b, &amp, ampl))) return rc;
  void* rmdef;
  if ((rc = bupload_bulk(b, &rmdef, A.rmdef))) return rc;
  {
    int nfull = 0;
    const std::vector<int> list = tile_list(A.tiles, fits32(nb, (size_t)ntiles * SLOTS,
                                                      b->coef_mode == COEF_SHARED
                                                          ? shared_off
                                                          : cb * std::max(nd, b->nelast) * NP,
                                                      rs), nfull);
    int* tl = nullptr;
    if ((rc = bupload(b, &tl, list))) return rc;
    X.tlist = tl;
    X.nfull = nfull;
    X.npart = ntiles - nfull;
  }
  X.T.ntiles = ntiles;
  X.T.per_xcd = (ntiles + 7) / 8;
  X.T.tiles = tiles;
  X.T.grids = grids;
  X.T.spec_cnt = sc;
  X.T.spec_idx = si;
  X.T.NGp = nb;
  X.T.FPS = (size_t)ntiles * SLOTS;
  X.ndef = (int)A.defs.size();
  X.defs.node = defs;
  X.defs.ext = defx;
  X.part_off = po;
  X.dsrc = ds;
  X.dabso = dab;
  X.rmdef = rmdef;
  X.ampl = amp;
  X.ampl_nt = ampl_nt;
  // bulk arrays: fields, inverse mass, coefficients, partial forces
  char *d, *v, *a, *rm0, *co, *ge = nullptr, *fp;
  if ((rc = balloc(b, &d, nd * nb * rs))) return rc;
  if ((rc = balloc(b, &v, nd * nb * rs))) return rc;
  if ((rc = balloc(b, &a, nd * nb * rs))) return rc;
  if ((rc = balloc(b, &rm0, nb * rs))) return rc;
  if ((rc = balloc(b, &fp, (size_t)nd * X.T.FPS * rs))) return rc;
  const size_t co_per = b->coef_mode == COEF_SHARED ? 0
                        : (b->coef_mode == COEF_COMPACT ? (size_t)nd * NP : (size_t)b->nelast * NP);
  const size_t co_n = b->coef_mode == COEF_SHARED ? shared_off : cb * co_per;
  if ((rc = balloc(b, &co, co_n * rs))) return rc;
  if (b->coef_mode == COEF_COMPACT && (rc = balloc(b, &ge, cb * 3 * rs))) return rc;
  size_t cofs = 0, gofs = 0;
  for (int m = 0; m < n; ++m) {
    sem2d_solver* h = b->mem[m];
    const size_t NS = h->st.NS, o = b->node_base[m];
    for (int c = 0; c < nd; ++c) {
      BHIPCHK(b, hipMemcpyAsync(d + (c * nb + o) * rs, (char*)h->d + c * NS * rs, NS * rs, hipMemcpyDeviceToDevice, b->stream));
      BHIPCHK(b, hipMemcpyAsync(v + (c * nb + o) * rs, (char*)h->v + c * NS * rs, NS * rs, hipMemcpyDeviceToDevice, b->stream));
      BHIPCHK(b, hipMemcpyAsync(a + (c * nb + o) * rs, (char*)h->a + c * NS * rs, NS * rs, hipMemcpyDeviceToDevice, b->stream));
    }
    BHIPCHK(b, hipMemcpyAsync(rm0 + o * rs, h->st.rmass0, NS * rs, hipMemcpyDeviceToDevice, b->stream));
    const size_t ncf = b->coef_mode == COEF_SHARED ? (size_t)h->nelast * NP : (size_t)h->nelem * co_per;
    BHIPCHK(b, hipMemcpyAsync(co + cofs * rs, h->st.coefs, ncf * rs, hipMemcpyDeviceToDevice, b->stream));
    cofs += ncf;
    if (ge) {
      BHIPCHK(b, hipMemcpyAsync(ge + gofs * rs, h->st.geom, (size_t)h->nelem * 3 * rs, hipMemcpyDeviceToDevice, b->stream));
      gofs += (size_t)h->nelem * 3;
    }
  }
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  X.d = d;
  X.v = v;
  X.a = a;
  X.rmass0 = rm0;
  X.coefs = co;
  X.geom = ge;
  X.fpart = fp;
  if ((rc = balloc(b, &b->it_dev, 1))) return rc;
  if ((rc = balloc(b, &b->flags, 4))) return rc;
  const int zero4[4] = {0, 0, 0, 0};
  BHIPCHK(b, hipMemcpy(b->flags, zero4, sizeof zero4, hipMemcpyHostToDevice));
  BHIPCHK(b, hipMemcpy(b->it_dev, zero4, sizeof(int), hipMemcpyHostToDevice));
  X.itp = b->it_dev;
  if (h0->flt.on) {
    if ((rc = gather_fault(b))) return rc;
    std::vector<double> r1, r2;
    for (int c = 0; c < nd; ++c) {
      r1.insert(r1.end(), frm_c[c].begin(), frm_c[c].end());
      r2.insert(r2.end(), frm2_c[c].begin(), frm2_c[c].end());
    }
    double *p1, *p2;
    if ((rc = bupload(b, &p1, r1))) return rc;
    if ((rc = bupload(b, &p2, r2))) return rc;
    b->X.F.rmn1 = p1;
    b->X.F.rmn2 = p2;
  }
  // receivers: members with stations, member after member
  std::vector<int> rn;
  b->rec_base.assign(n, 0);
  for (int m = 0; m < n; ++m) {
    sem2d_solver* h = b->mem[m];
    b->rec_base[m] = (int)rn.size();
    if (!h->nrec) continue;
    if (h->rec_w) return berr(b, 3, "member %d: receivers inside elements (AtNode=F) are not batched", m);
    if (!rn.empty() && (h->isamp != b->isamp || h->field != b->field))
      return berr(b, 3, "member %d: receiver sampling and field must match", m);
    b->isamp = h->isamp;
    b->field = h->field;
    b->nsamp = std::max(b->nsamp, h->nsamp);
    std::vector<int> t(h->nrec);
    BHIPCHK(b, hipMemcpy(t.data(), h->rec_node, h->nrec * sizeof(int), hipMemcpyDeviceToHost));
    for (int g : t) rn.push_back(g + (int)b->node_base[m]);
  }
  b->nrec = (int)rn.size();
  if ((rc = bupload(b, &b->rec_node, rn))) return rc;
  if ((rc = balloc(b, &b->seis, (size_t)std::max(b->nsamp, 1) * nd * std::max(b->nrec, 1)))) return rc;
  // default launch configuration (sem2d_batch_tune measures the alternatives)
  batch_set_config(b, 256, ntiles >= 8 * b->ncu ? (b->fp32 ? 3 : 2) : 0);
  // records of step 0, then the counter moves to step 1
  if (b->fp32) launch_batch_record<float>(b);
  else launch_batch_record<double>(b);
  k_tick<<<1, 64, 0, b->stream>>>(b->it_dev);
  BHIPCHK(b, hipGetLastError());
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  return 0;
}

int sem2d_batch_step(sem2d_batch* b, int nsteps) {
  if (!b) return 1;
  if (nsteps < 0) return berr(b, 3, "batch_step: negative nsteps");
  return batch_run(b, nsteps, true);
}

int sem2d_batch_tune(sem2d_batch* b, int trial_steps, int* choice) {
  if (!b) return 1;
  if (trial_steps < 1) return berr(b, 3, "batch_tune: trial_steps must be >= 1");
  // save the state the trials advance
  const size_t rs = b->fp32 ? sizeof(float) : sizeof(double);
  const size_t nf = (size_t)b->ndof * b->fstride * rs;
  std::vector<void*> saved;
  auto keep = [&](const void* src, size_t bytes) -> int {
    void* q = nullptr;
    BHIPCHK(b, hipMalloc(&q, std::max<size_t>(bytes, 1)));
    saved.push_back(q);
    BHIPCHK(b, hipMemcpyAsync(q, src, bytes, hipMemcpyDeviceToDevice, b->stream));
    return 0;
  };
  struct Blk { void* p; size_t bytes; };
  std::vector<Blk> blks = {{b->X.d, nf}, {b->X.v, nf}, {b->X.a, nf}, {b->it_dev, sizeof(int)}, {b->flags, 4 * sizeof(int)}};
  if (b->X.fault_on) {
    const FaultDev& F = b->X.F;
    const size_t np = F.np, ndp = (size_t)F.ndof * np;
    blks.push_back({F.theta, np * 8});
    blks.push_back({F.D, ndp * 8});
    blks.push_back({F.V, ndp * 8});
    blks.push_back({F.T, 2 * np * 8});
    blks.push_back({F.Tstick, 2 * np * 8});
    blks.push_back({F.MU, np * 8});
    blks.push_back({F.sigma, np * 8});
  }
  int rc = 0;
  for (auto& k : blks)
    if ((rc = keep(k.p, k.bytes))) break;
  const int it0 = b->it;
  const int cand[][2] = {{256, 3}, {512, 2}, {256, 2}, {512, 1}, {256, 1}, {256, 0}};
  const int ncand = sizeof(cand) / sizeof(cand[0]);
  double best = 1e30;
  int ibest = 0;
  for (int q = 0; q < ncand && !rc; ++q) {
    batch_set_config(b, cand[q][0], cand[q][1]);
    rc = batch_run(b, 1, false);  // warm-up
    if (!rc) rc = batch_run(b, trial_steps, false);
    if (!rc && b->last_ms < best) {
      best = b->last_ms;
      ibest = q;
    }
    // restore
    for (size_t k = 0; k < blks.size() && !rc; ++k)
      if (hipMemcpyAsync(blks[k].p, saved[k], blks[k].bytes, hipMemcpyDeviceToDevice, b->stream) != hipSuccess)
        rc = berr(b, 10, "batch_tune: restore failed");
    b->it = it0;
  }
  hipStreamSynchronize(b->stream);
  for (void* q : saved) hipFree(q);
  if (rc) return rc;
  batch_set_config(b, cand[ibest][0], cand[ibest][1]);
  if (choice) *choice = ibest;
  return 0;
}

int sem2d_batch_set_config(sem2d_batch* b, int threads, int wg_per_cu) {
  if (!b) return 1;
  if ((threads != 256 && threads != 512) || wg_per_cu < 0 || wg_per_cu > 4)
    return berr(b, 3, "batch_set_config: threads must be 256 or 512 and 0 <= wg_per_cu <= 4");
  batch_set_config(b, threads, wg_per_cu);
  return 0;
}

int sem2d_batch_config(sem2d_batch* b, int* threads, int* wgs, int* ntiles) {
  if (!b) return 1;
  if (threads) *threads = b->X.pipe_wgs > 0 ? b->X.pipe_threads : TILE_THREADS;
  if (wgs) *wgs = b->X.pipe_wgs > 0 ? b->X.pipe_wgs : 8 * b->X.T.per_xcd;
  if (ntiles) *ntiles = b->X.T.ntiles;
  return 0;
}

int sem2d_batch_fetch_fields(sem2d_batch* b, int m, double* d, double* v, double* a) {
  if (!b) return 1;
  if (m < 0 || m >= b->n) return berr(b, 3, "batch_fetch_fields: bad member %d", m);
  sem2d_solver* h = b->mem[m];
  const size_t NS = h->st.NS, o = b->node_base[m], rs = b->fp32 ? 4 : 8;
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  double* outs[3] = {d, v, a};
  void* srcs[3] = {b->X.d, b->X.v, b->X.a};
  std::vector<char> raw(NS * rs);
  for (int q = 0; q < 3; ++q) {
    if (!outs[q]) continue;
    for (int c = 0; c < b->ndof; ++c) {
      BHIPCHK(b, hipMemcpy(raw.data(), (char*)srcs[q] + (c * b->fstride + o) * rs, NS * rs, hipMemcpyDeviceToHost));
      for (int k = 0; k < h->npoin; ++k) {
        const size_t g = h->st.gmap[k];
        outs[q][(size_t)c * h->npoin + k] = b->fp32 ? (double)((float*)raw.data())[g] : ((double*)raw.data())[g];
      }
    }
  }
  return 0;
}

int sem2d_batch_fetch_seismo(sem2d_batch* b, int m, double* out, int* taken) {
  if (!b) return 1;
  if (m < 0 || m >= b->n) return berr(b, 3, "batch_fetch_seismo: bad member %d", m);
  sem2d_solver* h = b->mem[m];
  const int tk = h->nrec ? std::min(rec_taken(b), h->nsamp) : 0;
  if (out && tk) {
    std::vector<double> all;
    int rc = bdown(b, all, b->seis, (size_t)tk * b->ndof * b->nrec);
    if (rc) return rc;
    for (int s = 0; s < tk; ++s)
      for (int c = 0; c < b->ndof; ++c)
        for (int r = 0; r < h->nrec; ++r)
          out[((size_t)s * b->ndof + c) * h->nrec + r] = all[((size_t)s * b->ndof + c) * b->nrec + b->rec_base[m] + r];
  }
  if (taken) *taken = tk;
  return 0;
}

int sem2d_batch_fetch_fault(sem2d_batch* b, int m, double* out, int* taken) {
  if (!b) return 1;
  if (m < 0 || m >= b->n) return berr(b, 3, "batch_fetch_fault: bad member %d", m);
  if (!b->X.fault_on) return berr(b, 4, "batch_fetch_fault: no fault");
  const FaultHost& H = b->mem[m]->flt;
  const int tk = std::min(flt_taken(b), H.nsamp);
  if (out && tk) {
    std::vector<double> all;
    int rc = bdown(b, all, b->frec, (size_t)tk * 6 * b->fnout);
    if (rc) return rc;
    for (int s = 0; s < tk; ++s)
      for (int q = 0; q < 6; ++q)
        for (int o = 0; o < H.nout; ++o)
          out[((size_t)s * 6 + q) * H.nout + o] = all[((size_t)s * 6 + q) * b->fnout + b->fout_base[m] + o];
  }
  if (taken) *taken = tk;
  return 0;
}

int sem2d_batch_fetch_fault_state(sem2d_batch* b, int m, double* D, double* V, double* T, double* MU, double* Ts) {
  if (!b) return 1;
  if (m < 0 || m >= b->n) return berr(b, 3, "batch_fetch_fault_state: bad member %d", m);
  if (!b->X.fault_on) return berr(b, 4, "batch_fetch_fault_state: no fault");
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  const FaultDev& F = b->X.F;
  const int nm = b->mem[m]->flt.dev.np, o = b->fault_base[m], np = F.np;
  auto rows = [&](double* dst, const double* src, int nr) -> int {
    if (!dst) return 0;
    for (int r = 0; r < nr; ++r)
      BHIPCHK(b, hipMemcpy(dst + (size_t)r * nm, src + (size_t)r * np + o, nm * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
  };
  int rc;
  if ((rc = rows(D, F.D, F.ndof))) return rc;
  if ((rc = rows(V, F.V, F.ndof))) return rc;
  if ((rc = rows(T, F.T, 2))) return rc;
  if ((rc = rows(MU, F.MU, 1))) return rc;
  if ((rc = rows(Ts, F.Tstick, 2))) return rc;
  return 0;
}

int sem2d_batch_fetch_potency(sem2d_batch* b, int m, double* out, int* nrows) {
  if (!b) return 1;
  if (m < 0 || m >= b->n) return berr(b, 3, "batch_fetch_potency: bad member %d", m);
  if (!b->X.fault_on) return berr(b, 4, "batch_fetch_potency: no fault");
  const int rows = b->pot_shift + 1 + b->it;
  const size_t nc = 2 * (size_t)(b->ndof + 1);
  if (out && rows)
    BHIPCHK(b, hipMemcpy(out, b->pot + (size_t)m * b->pot_cap * nc, (size_t)rows * nc * sizeof(double),
                         hipMemcpyDeviceToHost));
  if (nrows) *nrows = rows;
  return 0;
}

int sem2d_batch_get_step(sem2d_batch* b, int* it) {
  if (!b || !it) return 1;
  *it = b->it;
  return 0;
}

int sem2d_batch_last_step_ms(sem2d_batch* b, double* ms) {
  if (!b || !ms) return 1;
  *ms = b->last_ms;
  return 0;
}

const char* sem2d_batch_last_error(sem2d_batch* b) { return b ? b->err.c_str() : "null handle"; }

void sem2d_batch_destroy(sem2d_batch* b) {
  if (!b) return;
  if (b->stream) hipStreamSynchronize(b->stream);
  batch_drop_graph(b);
  for (void* p : b->allocs) hipFree(p);
  if (b->pot) hipFree(b->pot);
  if (b->ev0) hipEventDestroy(b->ev0);
  if (b->ev1) hipEventDestroy(b->ev1);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
}

}  // extern "C"

// ===========================================================================
// Adjoint of the time loop with sqrt(N) checkpointing (structured P-SV grids,
// compact coefficients, beta = 0, no fault).
//
// forward:  from the zero state, nsteps steps; the state before every segment of
//           seg steps is kept (checkpoints); seismograms are recorded as usual.
// backward: for each segment, last first: the segment is replayed from its
//           checkpoint into seg+1 state slots, then the adjoint steps run from the
//           segment's last step down to its first, injecting dJ/d(seismogram) at the
//           sample steps and accumulating the sensitivities:
//             glam, gmu [cell]   dJ/d(lambda), dJ/d(mu) of each element (summed over
//                                its GLL points)
//             mbar [node]        dJ/dM (M the assembled mass, same for both components)
//             gabs [c][node]     dJ/dC_c at absorbing-boundary nodes (C the
//                                Clayton-Engquist coefficient, incl. its share c3*C
//                                of the modified mass)
// The host turns these into gradients with respect to per-element rho, lambda, mu
// (the chain rule through M = sum rho w J and C = sum rho c w |J1D|).
// ===========================================================================
struct AdjState {
  int nsteps = 0, seg = 0, nck = 0;
  int wgs = 256;  // workgroups of the adjoint tile kernel: one per CU, a multiple of 8
  // history mode: the displacement after every step (FP32 bulk: and the acceleration)
  // kept in device memory, so that the backward sweep needs no replay
  bool hist = false;
  char *H = nullptr, *AH = nullptr, *aN = nullptr;  // [nsteps+1] states; a after the last step
  void* hblock = nullptr;  // the history's device block (from or back to the history cache)
  size_t hbytes = 0;
  // injection plan of dJ/d(seismogram): distinct nodes, their contributions
  int nu = 0;
  int *unode = nullptr, *uoff = nullptr, *urec = nullptr;
  double* uw = nullptr;
  // the injection of step n-1 done by the deferred kernel of step n: per deferred node the
  // injection node it finishes (or -1), and the injection nodes that are not deferred
  int *udef = nullptr, *uext = nullptr;
  int next = 0;
  size_t state_bytes = 0;  // d, v, a of one state
  std::vector<char*> ck, slot;
  char *ad = nullptr, *av = nullptr, *aa = nullptr, *rmass = nullptr;
  double *glam = nullptr, *gmu = nullptr, *mbar = nullptr, *gabs = nullptr, *gseis = nullptr;
  std::vector<void*> allocs;
  double fwd_ms = 0, bwd_ms = 0;
};

// One reusable history block per process.  Releasing a block of this size (a hundred
// GB) and mapping a new one costs seconds, while a sequence of gradients (one adjoint
// state each) needs the same block again and again: a freed state's history block is
// kept here and handed to the next history plan that fits in it, until
// sem2d_adjoint_release.  A block in use by a live state is not shared.
struct HistCache {
  std::mutex m;
  void* p = nullptr;
  size_t bytes = 0;
  int dev = -1;
};
HistCache& hist_cache() {
  static HistCache c;
  return c;
}

// a block of at least `bytes` on the current device: the cached one if it fits, else a new
// one (the cached block released first); nullptr if the device has no room
void* hist_acquire(size_t bytes, size_t* got) {
  HistCache& c = hist_cache();
  std::lock_guard<std::mutex> lk(c.m);
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (c.p && c.dev == dev && c.bytes >= bytes) {
    void* p = c.p;
    *got = c.bytes;
    c.p = nullptr;
    c.bytes = 0;
    return p;
  }
  if (c.p) {
    (void)hipFree(c.p);
    c.p = nullptr;
    c.bytes = 0;
  }
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  *got = bytes;
  return p;
}

// size of the cached block usable on the current device (0: none)
size_t hist_cached_bytes() {
  HistCache& c = hist_cache();
  std::lock_guard<std::mutex> lk(c.m);
  int dev = 0;
  (void)hipGetDevice(&dev);
  return c.p && c.dev == dev ? c.bytes : 0;
}

void hist_return(void* p, size_t bytes) {
  HistCache& c = hist_cache();
  std::lock_guard<std::mutex> lk(c.m);
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (c.p && c.bytes >= bytes) {  // keep the larger block
    (void)hipFree(p);
    return;
  }
  if (c.p) (void)hipFree(c.p);
  c.p = p;
  c.bytes = bytes;
  c.dev = dev;
}

void adj_free(AdjState* A) {
  for (void* p : A->allocs) hipFree(p);
  if (A->hblock) hist_return(A->hblock, A->hbytes);
  delete A;
}

namespace {

int adj_alloc(sem2d_solver* h, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 1));
  if (e != hipSuccess) return set_err(h, 11, "adjoint: hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
  h->adj->allocs.push_back(*p);
  return 0;
}

// Inverse mass per component in storage order for the adjoint: the single-component
// array of the tile kernels (k_adj_rmass), then the deferred nodes' own values
// (k_adj_rmass_def).
template <typename R>
__global__ void k_adj_rmass(size_t GS, int ndof, const R* __restrict__ rm0, R* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= GS) return;
  const R m = rm0[i];
  for (int c = 0; c < ndof; ++c) out[(size_t)c * GS + i] = m;
}
template <typename R>
__global__ void k_adj_rmass_def(int ndef, int ndof, DefTable defs, const R* __restrict__ rmdef, size_t GS,
                                R* __restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ndef) return;
  const size_t g = defs.node[q].g;
  for (int c = 0; c < ndof; ++c) out[(size_t)c * GS + g] = rmdef[(size_t)c * ndef + q];
}

// copy a whole state (d, v, a) between two state buffers or the solver's fields
int copy_state(sem2d_solver* h, char* dst_d, char* dst_v, char* dst_a, const char* d, const char* v, const char* a) {
  const size_t n = (size_t)h->ndof * h->fstride * rsize(h);
  HIPCHK(h, hipMemcpyAsync(dst_d, d, n, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(dst_v, v, n, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(dst_a, a, n, hipMemcpyDeviceToDevice, h->stream));
  return 0;
}

inline char* st_d(sem2d_solver* h, char* s) { (void)h; return s; }
inline char* st_v(sem2d_solver* h, char* s) { return s + (size_t)h->ndof * h->fstride * rsize(h); }
inline char* st_a(sem2d_solver* h, char* s) { return s + 2 * (size_t)h->ndof * h->fstride * rsize(h); }

// One adjoint tile step: the full tiles of the launch list with the constant-shape
// variant (every array it addresses with 32-bit offsets below 4 GiB, the double
// mass-sensitivity array included), the others with the general one.  FP64 fields with
// a displacement history (the acceleration formed from d_{n-1}, d_n, d_{n+1}):
// persistent pipelined kernels of one 512-thread workgroup per CU (their LDS allows no
// more; 162 against 182 us per 400x400-element step).  Otherwise one 256-thread
// workgroup per tile (stored acceleration, FP64: 139 against 156 us; FP32 fields, four
// workgroups per CU: 91 against 114 us).
template <typename R>
void launch_adj_tile(sem2d_solver* h, const StepDev& X, const R* dfw, const R* afw, const R* dp, const R* dn,
                     R dt, R c1, R c2, R c3, double idt2) {
  AdjState& A = *h->adj;
  StructState& S = h->st;
  const int nf = (h->fstride * sizeof(double) < (size_t(1) << 32)) ? S.nfull : 0;
  const int np = S.ntiles - nf;
  if (sizeof(R) == 4 || afw) {
    if (nf > 0)
      k_adj_tile<R, true><<<8 * ((nf + 7) / 8), TILE_THREADS, 0, h->stream>>>(
          X.T, S.tlist_dev, nf, (R*)A.ad, (R*)A.av, (R*)A.aa, (const R*)A.rmass, dfw, afw, (const R*)S.coefs,
          (const R*)S.geom, (R*)S.fpart, A.glam, A.gmu, A.mbar, dt, c1, c2, c3, dp, dn, idt2);
    if (np > 0)
      k_adj_tile<R, false><<<8 * ((np + 7) / 8), TILE_THREADS, 0, h->stream>>>(
          X.T, S.tlist_dev + nf, np, (R*)A.ad, (R*)A.av, (R*)A.aa, (const R*)A.rmass, dfw, afw,
          (const R*)S.coefs, (const R*)S.geom, (R*)S.fpart, A.glam, A.gmu, A.mbar, dt, c1, c2, c3, dp, dn, idt2);
  } else if constexpr (sizeof(R) == 8) {
    auto wgs = [&](int n) { return std::min(A.wgs, (n + 7) / 8 * 8); };
    if (nf > 0)
      k_adj_pipe<R, true><<<wgs(nf), AP_NT, 0, h->stream>>>(
          X.T, S.tlist_dev, nf, X.T.tiles, X.T.grids, X.T.spec_cnt, (R*)A.ad, (R*)A.av, (R*)A.aa,
          (const R*)A.rmass, dfw, afw, (const R*)S.coefs, (const R*)S.geom, (R*)S.fpart, A.glam, A.gmu, A.mbar,
          dt, c1, c2, c3, dp, dn, idt2);
    if (np > 0)
      k_adj_pipe<R, false><<<wgs(np), AP_NT, 0, h->stream>>>(
          X.T, S.tlist_dev + nf, np, X.T.tiles, X.T.grids, X.T.spec_cnt, (R*)A.ad, (R*)A.av, (R*)A.aa,
          (const R*)A.rmass, dfw, afw, (const R*)S.coefs, (const R*)S.geom, (R*)S.fpart, A.glam, A.gmu, A.mbar,
          dt, c1, c2, c3, dp, dn, idt2);
  }
}

template <typename R>
int adj_backward(sem2d_solver* h, const double* gseis_host) {
  AdjState& A = *h->adj;
  StructState& S = h->st;
  const size_t GS = h->fstride, nf = (size_t)h->ndof * GS;
  const StepCoefs<R> K = step_coefs<R>(h->dt, h->beta, h->gamma, h->leapfrog);
  const R dt = K.dt, c1 = K.c1, c2 = K.c2, c3 = K.c3, c4 = K.c4;
  const size_t ns = (size_t)h->nsamp * h->ndof * h->nrec;
  if (h->nrec && gseis_host)
    HIPCHK(h, hipMemcpyAsync(A.gseis, gseis_host, ns * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemsetAsync(A.ad, 0, nf * sizeof(R), h->stream));
  HIPCHK(h, hipMemsetAsync(A.av, 0, nf * sizeof(R), h->stream));
  HIPCHK(h, hipMemsetAsync(A.aa, 0, nf * sizeof(R), h->stream));
  HIPCHK(h, hipMemsetAsync(A.glam, 0, (size_t)h->nelem * sizeof(double), h->stream));
  HIPCHK(h, hipMemsetAsync(A.gmu, 0, (size_t)h->nelem * sizeof(double), h->stream));
  HIPCHK(h, hipMemsetAsync(A.mbar, 0, GS * sizeof(double), h->stream));
  HIPCHK(h, hipMemsetAsync(A.gabs, 0, (size_t)h->ndof * std::max(S.ndef, 1) * sizeof(double), h->stream));
  StepDev X = step_dev(h);
  R* inj = (R*)(h->field == 0 ? A.ad : (h->field == 1 ? A.av : A.aa));
  auto sampled = [&](int n) { return h->nrec && n >= 1 && n % h->isamp == 0 && n / h->isamp < h->nsamp; };
  auto gs_of = [&](int n) { return A.gseis + (size_t)(n / h->isamp) * h->ndof * h->nrec; };
  // the deferred kernel of step n also adds the sample of step n-1 (after its own
  // updates, before the tile kernel of step n-1 reads the field); the first step's
  // sample gets its own launch
  auto plan_for = [&](int n) {
    InjectPlan P;
    if (!sampled(n - 1)) return P;
    P.nu = A.nu;
    P.nrec = h->nrec;
    P.unode = A.unode;
    P.off = A.uoff;
    P.rec = A.urec;
    P.w = A.uw;
    P.g = gs_of(n - 1);
    P.fld = inj;
    P.udef = A.udef;
    P.uext = A.uext;
    P.next = A.next;
    return P;
  };
  auto deferred_grid = [&](const InjectPlan& P) { return (S.ndef + (P.g ? P.next : 0) + 255) / 256; };
  HIPCHK(h, hipEventRecord(h->ev0, h->stream));
  if (sampled(A.nsteps))
    k_adj_inject<R><<<grid_for(A.nu, 64), 64, 0, h->stream>>>(A.nu, A.unode, A.uoff, A.urec, A.uw, h->nrec,
                                                             gs_of(A.nsteps), inj, GS);
  if (A.hist) {
    // no replay: the forward displacements are all in memory
    const size_t nfb = nf * sizeof(R);
    const double idt = 1.0 / h->dt, idt2 = idt * idt;
    const int N = A.nsteps;
    for (int n = N; n >= 1; --n) {
      const R* dc = (const R*)(A.H + (size_t)n * nfb);
      const R* dp = (const R*)(A.H + (size_t)(n - 1) * nfb);
      const R* dn = n < N ? (const R*)(A.H + (size_t)(n + 1) * nfb) : nullptr;
      const R* af = A.AH ? (const R*)(A.AH + (size_t)n * nfb) : (n == N ? (const R*)A.aN : nullptr);
      launch_adj_tile<R>(h, X, dc, af, dp, dn, dt, c1, c2, c3, idt2);
      const InjectPlan P = plan_for(n);
      k_adj_deferred<R><<<deferred_grid(P), 256, 0, h->stream>>>(
          S.ndef, S.defs, S.part_off, (const R*)S.fpart, S.FPS, S.dabso, (const R*)S.rmdef, (R*)A.ad, (R*)A.av,
          (R*)A.aa, GS, af, nullptr, nullptr, A.mbar, A.gabs, dt, c1, c2, c3, dc, dp, dn, idt, idt2, P);
    }
  }
  for (int k = A.hist ? -1 : A.nck - 1; k >= 0; --k) {
    const int s0 = k * A.seg, s1 = std::min(A.nsteps, s0 + A.seg), L = s1 - s0;
    // replay: slot[j] = state after step s0 + j
    char* s = A.slot[0];
    int rc = copy_state(h, st_d(h, s), st_v(h, s), st_a(h, s), st_d(h, A.ck[k]), st_v(h, A.ck[k]), st_a(h, A.ck[k]));
    if (rc) return rc;
    for (int j = 1; j <= L; ++j) {  // out of place: slot[j-1] -> slot[j]
      char* p = A.slot[j - 1];
      char* q = A.slot[j];
      X.din = st_d(h, p);
      X.vin = st_v(h, p);
      X.ain = st_a(h, p);
      X.d = st_d(h, q);
      X.v = st_v(h, q);
      X.a = st_a(h, q);
      launch_structured_step<R>(X, s0 + j, dt, c1, c2, c3, c4);
    }
    X.din = X.vin = X.ain = nullptr;
    // adjoint steps n = s1 .. s0+1
    for (int n = s1; n > s0; --n) {
      const int j = n - s0;
      const R* dfw = (const R*)st_d(h, A.slot[j]);
      const R* afw = (const R*)st_a(h, A.slot[j]);
      launch_adj_tile<R>(h, X, dfw, afw, nullptr, nullptr, dt, c1, c2, c3, 0.0);
      const InjectPlan P = plan_for(n);  // the sample of step n-1 (the replay leaves the adjoint alone)
      k_adj_deferred<R><<<deferred_grid(P), 256, 0, h->stream>>>(
          S.ndef, S.defs, S.part_off, (const R*)S.fpart, S.FPS, S.dabso, (const R*)S.rmdef, (R*)A.ad, (R*)A.av,
          (R*)A.aa, GS, afw, (const R*)st_v(h, A.slot[j - 1]), (const R*)st_a(h, A.slot[j - 1]), A.mbar, A.gabs,
          dt, c1, c2, c3, nullptr, nullptr, nullptr, 0.0, 0.0, P);
    }
  }
  HIPCHK(h, hipEventRecord(h->ev1, h->stream));
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  A.bwd_ms = ms;
  return 0;
}

// Forward sweep of the history mode.  Step n reads the displacement of history slot n-1
// and writes it into slot n (out of place, with v, and a unless it has a history, in
// place), so the history needs no copies; for FP32 fields the acceleration goes through
// its own history the same way.  The receivers read the current slots; the solver's own
// arrays receive the final state, and the acceleration after the last step is kept.
template <typename R>
int adj_forward_hist(sem2d_solver* h) {
  AdjState& A = *h->adj;
  const size_t nfb = (size_t)h->ndof * h->fstride * sizeof(R);
  const StepCoefs<R> K = step_coefs<R>(h->dt, h->beta, h->gamma, h->leapfrog);
  const R dt = K.dt, c1 = K.c1, c2 = K.c2, c3 = K.c3, c4 = K.c4;
  int rc = sync_sources(h);
  if (rc) return rc;
  StepDev X = step_dev(h);
  void *d0 = h->d, *a0 = h->a;
  auto Hn = [&](int n) { return (void*)(A.H + (size_t)n * nfb); };
  auto AHn = [&](int n) { return (void*)(A.AH + (size_t)n * nfb); };
  HIPCHK(h, hipMemsetAsync(A.H, 0, nfb, h->stream));
  if (A.AH) HIPCHK(h, hipMemsetAsync(A.AH, 0, nfb, h->stream));
  HIPCHK(h, hipEventRecord(h->ev0, h->stream));
  for (int n = 1; n <= A.nsteps; ++n) {
    X.din = Hn(n - 1);
    X.vin = nullptr;  // in place
    X.ain = A.AH ? AHn(n - 1) : nullptr;
    X.d = Hn(n);
    X.a = A.AH ? AHn(n) : a0;
    launch_structured_step<R>(X, n, dt, c1, c2, c3, c4);
    h->it = n;
    h->d = X.d;
    h->a = X.a;
    rc = record_receivers<R>(h);
    h->d = d0;
    h->a = a0;
    if (rc) return rc;
  }
  HIPCHK(h, hipMemcpyAsync(d0, Hn(A.nsteps), nfb, hipMemcpyDeviceToDevice, h->stream));
  if (A.AH) HIPCHK(h, hipMemcpyAsync(a0, AHn(A.nsteps), nfb, hipMemcpyDeviceToDevice, h->stream));
  if (A.aN) HIPCHK(h, hipMemcpyAsync(A.aN, a0, nfb, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipEventRecord(h->ev1, h->stream));
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  A.fwd_ms = ms;
  return 0;
}

}  // namespace

extern "C" {

int sem2d_adjoint_init(sem2d_solver* h, int nsteps, int seg) {
  if (!h) return 1;
  if (!h->st.on || h->ndof != 2 || h->coef_mode != COEF_COMPACT || h->nelast != 6)
    return set_err(h, 4, "adjoint: needs a structured P-SV grid with compact (lambda, mu) coefficients");
  if (h->beta != 0.0 || h->gamma != 0.5 || h->leapfrog)
    return set_err(h, 4, "adjoint: needs the explicit Newmark scheme (beta = 0, gamma = 1/2)");
  if (h->flt.on) return set_err(h, 4, "adjoint: faults are not differentiated");
  if (h->st.eta || h->st.b25) return set_err(h, 4, "adjoint: Kelvin-Voigt and 2.5D terms are not differentiated");
  for (const AbsoHost& B : h->abso_host)
    if (!B.senode.empty()) return set_err(h, 4, "adjoint: Stacey boundaries are not differentiated");
  if (nsteps < 1) return set_err(h, 3, "adjoint_init: nsteps must be >= 1");
  const bool autoplan = seg <= 0;
  if (seg <= 0) seg = (int)std::ceil(std::sqrt((double)nsteps));
  int rc;
  if (!h->st.ready && (rc = sync_sources(h))) return rc;
  if (!h->st.ready && (rc = prepare_structured(h))) return rc;
  if (h->adj) adj_free(h->adj);
  if (h->flt.pot) hipFree(h->flt.pot);
  h->adj = new AdjState();
  AdjState& A = *h->adj;
  {
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    A.wgs = std::max(8, (ncu + 7) / 8 * 8);
  }
  A.nsteps = nsteps;
  A.seg = std::min(seg, nsteps);
  A.nck = (nsteps + A.seg - 1) / A.seg;
  const size_t nf = (size_t)h->ndof * h->fstride * rsize(h);
  A.state_bytes = 3 * nf;
  if (autoplan) {
    // history mode when every step's displacement (FP32 fields: and acceleration) fits in
    // free device memory, next to the adjoint's own arrays and a 4 GiB reserve
    const size_t hbytes = (size_t)(nsteps + 1) * nf * (h->fp32 ? 2 : 1);
    size_t fr = 0, tot = 0;
    const size_t cached = hist_cached_bytes();  // reusable, so counted as free
    if (hipMemGetInfo(&fr, &tot) == hipSuccess &&
        hbytes + 9 * nf + ((size_t)4 << 30) < fr + (cached >= hbytes ? hbytes : cached)) {
      size_t got = 0;
      void* p = hist_acquire(hbytes, &got);
      if (p) {
        A.hblock = p;
        A.hbytes = got;
        A.H = (char*)p;
        A.hist = true;
        if (h->fp32) {
          A.AH = A.H + (size_t)(nsteps + 1) * nf;
        } else {
          if ((rc = adj_alloc(h, &p, nf))) return rc;
          A.aN = (char*)p;
        }
        A.seg = nsteps;
        A.nck = 1;
      }  // else: checkpoints
    }
  }
  for (int k = 0; !A.hist && k < A.nck; ++k) {
    void* p;
    if ((rc = adj_alloc(h, &p, A.state_bytes))) return rc;
    A.ck.push_back((char*)p);
  }
  for (int k = 0; !A.hist && k <= A.seg; ++k) {
    void* p;
    if ((rc = adj_alloc(h, &p, A.state_bytes))) return rc;
    A.slot.push_back((char*)p);
  }
  void* p;
  if ((rc = adj_alloc(h, &p, nf))) return rc;
  A.ad = (char*)p;
  if ((rc = adj_alloc(h, &p, nf))) return rc;
  A.av = (char*)p;
  if ((rc = adj_alloc(h, &p, nf))) return rc;
  A.aa = (char*)p;
  if ((rc = adj_alloc(h, &p, (size_t)h->nelem * sizeof(double)))) return rc;
  A.glam = (double*)p;
  if ((rc = adj_alloc(h, &p, (size_t)h->nelem * sizeof(double)))) return rc;
  A.gmu = (double*)p;
  if ((rc = adj_alloc(h, &p, h->fstride * sizeof(double)))) return rc;
  A.mbar = (double*)p;
  if ((rc = adj_alloc(h, &p, (size_t)h->ndof * std::max(h->st.ndef, 1) * sizeof(double)))) return rc;
  A.gabs = (double*)p;
  if ((rc = adj_alloc(h, &p, (size_t)std::max(h->nsamp, 1) * h->ndof * std::max(h->nrec, 1) * sizeof(double))))
    return rc;
  A.gseis = (double*)p;
  // injection plan: the receivers' nodes (or element nodes with weights), grouped by
  // node, contributions in receiver order
  if (h->nrec) {
    const int per = h->rec_w ? h->np : 1;
    std::vector<int> rn((size_t)h->nrec * per);
    std::vector<double> rw((size_t)h->nrec * per, 1.0);
    HIPCHK(h, hipMemcpy(rn.data(), h->rec_node, rn.size() * sizeof(int), hipMemcpyDeviceToHost));
    if (h->rec_w) HIPCHK(h, hipMemcpy(rw.data(), h->rec_w, rw.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::map<int, std::vector<std::pair<int, double>>> by;  // node -> (receiver, weight)
    for (int r = 0; r < h->nrec; ++r)
      for (int k = 0; k < per; ++k) by[rn[(size_t)r * per + k]].push_back({r, rw[(size_t)r * per + k]});
    std::vector<int> un, off(1, 0), rec;
    std::vector<double> w;
    for (auto& kv : by) {
      un.push_back(kv.first);
      for (auto& e : kv.second) {
        rec.push_back(e.first);
        w.push_back(e.second);
      }
      off.push_back((int)rec.size());
    }
    A.nu = (int)un.size();
    void* q;
    if ((rc = adj_alloc(h, &q, un.size() * sizeof(int)))) return rc;
    A.unode = (int*)q;
    if ((rc = adj_alloc(h, &q, off.size() * sizeof(int)))) return rc;
    A.uoff = (int*)q;
    if ((rc = adj_alloc(h, &q, rec.size() * sizeof(int)))) return rc;
    A.urec = (int*)q;
    if ((rc = adj_alloc(h, &q, w.size() * sizeof(double)))) return rc;
    A.uw = (double*)q;
    HIPCHK(h, hipMemcpy(A.unode, un.data(), un.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(A.uoff, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(A.urec, rec.data(), rec.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(A.uw, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice));
    // injection nodes that are deferred nodes (storage index of deferred node q:
    // gmap[def_ref[q]]) are finished by that deferred lane, the others by extra lanes
    const StructState& S = h->st;
    std::unordered_map<int, int> u_at;  // injection node -> u (a few entries)
    for (int u = 0; u < (int)un.size(); ++u) u_at[un[u]] = u;
    std::vector<int> udef(std::max(S.ndef, 1), -1), uext;
    std::vector<char> taken(un.size(), 0);
    for (int q = 0; q < S.ndef; ++q) {
      auto it = u_at.find(S.gmap[S.def_ref[q]]);
      if (it != u_at.end()) {
        udef[q] = it->second;
        taken[it->second] = 1;
      }
    }
    for (int u = 0; u < (int)un.size(); ++u)
      if (!taken[u]) uext.push_back(u);
    A.next = (int)uext.size();
    if ((rc = adj_alloc(h, &q, udef.size() * sizeof(int)))) return rc;
    A.udef = (int*)q;
    if ((rc = adj_alloc(h, &q, std::max<size_t>(uext.size(), 1) * sizeof(int)))) return rc;
    A.uext = (int*)q;
    HIPCHK(h, hipMemcpy(A.udef, udef.data(), udef.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!uext.empty())
      HIPCHK(h, hipMemcpy(A.uext, uext.data(), uext.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  // inverse mass of both components in storage order.  The tile kernels' single-
  // component array holds it at every node whose components are equal; the deferred
  // nodes' per-component array covers the others (absorbing-boundary nodes), so the
  // device assembles it from those two when every node with unequal components is a
  // deferred node (else from the host copy).
  if ((rc = adj_alloc(h, &p, nf))) return rc;
  A.rmass = (char*)p;
  const StructState& S = h->st;
  bool dev_ok = true;
  {
    std::vector<char> isdef(h->npoin, 0);
    for (int q = 0; q < S.ndef; ++q) isdef[S.def_ref[q]] = 1;
    const double* rh = h->rmass_host.data();
    for (int c = 1; c < h->ndof && dev_ok; ++c)
      for (int k = 0; k < h->npoin; ++k)
        if (rh[(size_t)c * h->npoin + k] != rh[k] && !isdef[k]) {
          dev_ok = false;
          break;
        }
  }
  if (dev_ok) {
    const size_t GS = h->fstride;
    if (h->fp32) {
      k_adj_rmass<float><<<grid_for(GS, 256), 256, 0, h->stream>>>(GS, h->ndof, (const float*)S.rmass0,
                                                                  (float*)A.rmass);
      if (S.ndef > 0)
        k_adj_rmass_def<float><<<grid_for(S.ndef, 256), 256, 0, h->stream>>>(
            S.ndef, h->ndof, S.defs, (const float*)S.rmdef, GS, (float*)A.rmass);
    } else {
      k_adj_rmass<double><<<grid_for(GS, 256), 256, 0, h->stream>>>(GS, h->ndof, (const double*)S.rmass0,
                                                                    (double*)A.rmass);
      if (S.ndef > 0)
        k_adj_rmass_def<double><<<grid_for(S.ndef, 256), 256, 0, h->stream>>>(
            S.ndef, h->ndof, S.defs, (const double*)S.rmdef, GS, (double*)A.rmass);
    }
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipStreamSynchronize(h->stream));
  } else {
    std::vector<double> rm((size_t)h->ndof * h->fstride, 0.0);
    for (int c = 0; c < h->ndof; ++c)
      for (int k = 0; k < h->npoin; ++k)
        rm[(size_t)c * h->fstride + h->st.gmap[k]] = h->rmass_host[(size_t)c * h->npoin + k];
    if ((rc = write_bulk(h, A.rmass, rm.data(), rm.size()))) return rc;
  }
  return 0;
}

int sem2d_adjoint_forward(sem2d_solver* h) {
  if (!h) return 1;
  if (!h->adj) return set_err(h, 4, "adjoint_forward: call sem2d_adjoint_init first");
  AdjState& A = *h->adj;
  const size_t nf = (size_t)h->ndof * h->fstride * rsize(h);
  HIPCHK(h, hipMemsetAsync(h->d, 0, nf, h->stream));
  HIPCHK(h, hipMemsetAsync(h->v, 0, nf, h->stream));
  HIPCHK(h, hipMemsetAsync(h->a, 0, nf, h->stream));
  h->it = 0;
  h->taken = 0;
  int rc = h->fp32 ? record_receivers<float>(h) : record_receivers<double>(h);
  if (rc) return rc;
  if (A.hist) return h->fp32 ? adj_forward_hist<float>(h) : adj_forward_hist<double>(h);
  double ms = 0;
  for (int k = 0; k < A.nck; ++k) {
    char* c = A.ck[k];
    if ((rc = copy_state(h, st_d(h, c), st_v(h, c), st_a(h, c), (char*)h->d, (char*)h->v, (char*)h->a))) return rc;
    const int L = std::min(A.seg, A.nsteps - k * A.seg);
    rc = h->fp32 ? do_steps<float>(h, L) : do_steps<double>(h, L);
    if (rc) return rc;
    ms += h->last_ms * L;
  }
  A.fwd_ms = ms;
  return 0;
}

int sem2d_adjoint_backward(sem2d_solver* h, const double* gseis) {
  if (!h) return 1;
  if (!h->adj) return set_err(h, 4, "adjoint_backward: call sem2d_adjoint_init first");
  if (h->it != h->adj->nsteps) return set_err(h, 4, "adjoint_backward: run sem2d_adjoint_forward first");
  return h->fp32 ? adj_backward<float>(h, gseis) : adj_backward<double>(h, gseis);
}

int sem2d_adjoint_fetch(sem2d_solver* h, double* glam, double* gmu, double* gmass, double* gabso) {
  if (!h) return 1;
  if (!h->adj) return set_err(h, 4, "adjoint_fetch: call sem2d_adjoint_init first");
  AdjState& A = *h->adj;
  StructState& S = h->st;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  std::vector<double> tmp;
  if (glam || gmu) {
    tmp.resize(h->nelem);
    for (int q = 0; q < 2; ++q) {
      double* out = q == 0 ? glam : gmu;
      if (!out) continue;
      HIPCHK(h, hipMemcpy(tmp.data(), q == 0 ? A.glam : A.gmu, h->nelem * sizeof(double), hipMemcpyDeviceToHost));
      for (int cell = 0; cell < h->nelem; ++cell) out[S.cell_elem[cell]] = tmp[cell];
    }
  }
  if (gmass) {
    tmp.resize(h->fstride);
    HIPCHK(h, hipMemcpy(tmp.data(), A.mbar, h->fstride * sizeof(double), hipMemcpyDeviceToHost));
    for (int k = 0; k < h->npoin; ++k) gmass[k] = tmp[S.gmap[k]];
  }
  if (gabso) {
    std::fill(gabso, gabso + (size_t)h->ndof * h->npoin, 0.0);
    tmp.resize((size_t)h->ndof * std::max(S.ndef, 1));
    HIPCHK(h, hipMemcpy(tmp.data(), A.gabs, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int c = 0; c < h->ndof; ++c)
      for (int q = 0; q < S.ndef; ++q) gabso[(size_t)c * h->npoin + S.def_ref[q]] = tmp[(size_t)c * S.ndef + q];
  }
  return 0;
}

int sem2d_adjoint_release(void) {
  HistCache& c = hist_cache();
  std::lock_guard<std::mutex> lk(c.m);
  if (c.p) {
    hipError_t e = hipFree(c.p);
    c.p = nullptr;
    c.bytes = 0;
    if (e != hipSuccess) return 10;
  }
  return 0;
}

int sem2d_adjoint_plan(sem2d_solver* h, int* history, int* seg) {
  if (!h) return 1;
  if (!h->adj) return set_err(h, 4, "adjoint_plan: call sem2d_adjoint_init first");
  if (history) *history = h->adj->hist ? 1 : 0;
  if (seg) *seg = h->adj->seg;
  return 0;
}

int sem2d_adjoint_timing(sem2d_solver* h, double* fwd_ms, double* bwd_ms) {
  if (!h) return 1;
  if (!h->adj) return set_err(h, 4, "adjoint_timing: call sem2d_adjoint_init first");
  if (fwd_ms) *fwd_ms = h->adj->fwd_ms;
  if (bwd_ms) *bwd_ms = h->adj->bwd_ms;
  return 0;
}

}  // extern "C"
