// Overset fringe interpolation on gfx950.
//
// Restates interpolation_tet (interpolation.f90:5-76): receivers are the
// points of a boundary mesh; donors are the elements with ELM_ID == src_id,
// taken in element index order; the first element that contains a point
// wins and the value of ug is interpolated there.  Tets use the reference's
// own test (barycentric coordinates from the inverse edge matrix, all >= 0,
// no tolerance), including its receiver offset bvec = xr - x_1 + dr, which
// leaves out the donor's first-node displacement (interpolation.f90:49-50).
// Hexes, which the reference skips (interpolation.f90:28), use local
// coordinates from a fixed number of Newton steps with the linhexshl shape
// functions, accepted in [-1 - 1e-10, 1 + 1e-10]^3 among the elements whose
// widened bounding box holds the point.  setBCs_overset's step 3
// (setBCs.f90:158-165) applies the values as Dirichlet data.
//
// One 64-lane wavefront per receiver scans the source elements in chunks of
// 64 in index order: each lane tests one element (a bounding-box test
// first), a ballot gives the hits and the lowest lane is the first element
// in index order, so the result is the reference's first-wins rule while
// every wave stops at its donor's chunk.  Arithmetic follows the reference's
// operation order with contraction off, so donors and values are the same
// bits as the CPU restatement.
#include "ofd_internal.h"

#include <mutex>

namespace ofd {

namespace {

constexpr int kHexNewton = 6;
constexpr double kHexTol = 1e-10;
constexpr double kBoxMargin = 1e-8;
constexpr int kTetGeo = 15;  // inverse edge matrix (9), first corner x (3) and d (3)
constexpr int kHexGeo = 24;  // the 8 corners in the current configuration

#define OFD_M3(A, i, j) (A)[(i) + 3 * (j)]

// get_inverse_3x3 (eval_shape.f90:85-112): cofactors, DetJ = 1/(1/det)
__device__ void inverse3x3(const double *A, double *Ai) {
#pragma clang fp contract(off)
  OFD_M3(Ai, 0, 0) = OFD_M3(A, 1, 1) * OFD_M3(A, 2, 2) - OFD_M3(A, 2, 1) * OFD_M3(A, 1, 2);
  OFD_M3(Ai, 0, 1) = OFD_M3(A, 2, 1) * OFD_M3(A, 0, 2) - OFD_M3(A, 0, 1) * OFD_M3(A, 2, 2);
  OFD_M3(Ai, 0, 2) = OFD_M3(A, 0, 1) * OFD_M3(A, 1, 2) - OFD_M3(A, 0, 2) * OFD_M3(A, 1, 1);
  const double tmp = 1.0 / (OFD_M3(Ai, 0, 0) * OFD_M3(A, 0, 0) +
                            OFD_M3(Ai, 0, 1) * OFD_M3(A, 1, 0) +
                            OFD_M3(Ai, 0, 2) * OFD_M3(A, 2, 0));
  OFD_M3(Ai, 0, 0) = OFD_M3(Ai, 0, 0) * tmp;
  OFD_M3(Ai, 0, 1) = OFD_M3(Ai, 0, 1) * tmp;
  OFD_M3(Ai, 0, 2) = OFD_M3(Ai, 0, 2) * tmp;
  OFD_M3(Ai, 1, 0) = (OFD_M3(A, 1, 2) * OFD_M3(A, 2, 0) - OFD_M3(A, 1, 0) * OFD_M3(A, 2, 2)) * tmp;
  OFD_M3(Ai, 1, 1) = (OFD_M3(A, 0, 0) * OFD_M3(A, 2, 2) - OFD_M3(A, 2, 0) * OFD_M3(A, 0, 2)) * tmp;
  OFD_M3(Ai, 1, 2) = (OFD_M3(A, 1, 0) * OFD_M3(A, 0, 2) - OFD_M3(A, 0, 0) * OFD_M3(A, 1, 2)) * tmp;
  OFD_M3(Ai, 2, 0) = (OFD_M3(A, 1, 0) * OFD_M3(A, 2, 1) - OFD_M3(A, 1, 1) * OFD_M3(A, 2, 0)) * tmp;
  OFD_M3(Ai, 2, 1) = (OFD_M3(A, 2, 0) * OFD_M3(A, 0, 1) - OFD_M3(A, 0, 0) * OFD_M3(A, 2, 1)) * tmp;
  OFD_M3(Ai, 2, 2) = (OFD_M3(A, 0, 0) * OFD_M3(A, 1, 1) - OFD_M3(A, 0, 1) * OFD_M3(A, 1, 0)) * tmp;
}

// linhexshl (eval_shape_FEM.f90:148-196): N(8) and dN/dxi (8,3) column-major
__device__ void hex_shape(const double xi[3], double *N, double *dN) {
#pragma clang fp contract(off)
  const double z = xi[0], t = xi[1], u = xi[2], e = 0.125;
  N[0] = (z - 1.0) * (t + 1.0) * (u - 1.0) * e;
  N[1] = -(z + 1.0) * (t + 1.0) * (u - 1.0) * e;
  N[2] = (z + 1.0) * (t + 1.0) * (u + 1.0) * e;
  N[3] = -(z - 1.0) * (t + 1.0) * (u + 1.0) * e;
  N[4] = -(z - 1.0) * (t - 1.0) * (u - 1.0) * e;
  N[5] = (z + 1.0) * (t - 1.0) * (u - 1.0) * e;
  N[6] = -(z + 1.0) * (t - 1.0) * (u + 1.0) * e;
  N[7] = (z - 1.0) * (t - 1.0) * (u + 1.0) * e;
  dN[0] = (t + 1.0) * (u - 1.0) * e;
  dN[1] = -(t + 1.0) * (u - 1.0) * e;
  dN[2] = (t + 1.0) * (u + 1.0) * e;
  dN[3] = -(t + 1.0) * (u + 1.0) * e;
  dN[4] = -(t - 1.0) * (u - 1.0) * e;
  dN[5] = (t - 1.0) * (u - 1.0) * e;
  dN[6] = -(t - 1.0) * (u + 1.0) * e;
  dN[7] = (t - 1.0) * (u + 1.0) * e;
  dN[8 + 0] = (z - 1.0) * (u - 1.0) * e;
  dN[8 + 1] = -(z + 1.0) * (u - 1.0) * e;
  dN[8 + 2] = (z + 1.0) * (u + 1.0) * e;
  dN[8 + 3] = -(z - 1.0) * (u + 1.0) * e;
  dN[8 + 4] = -(z - 1.0) * (u - 1.0) * e;
  dN[8 + 5] = (z + 1.0) * (u - 1.0) * e;
  dN[8 + 6] = -(z + 1.0) * (u + 1.0) * e;
  dN[8 + 7] = (z - 1.0) * (u + 1.0) * e;
  dN[16 + 0] = (z - 1.0) * (t + 1.0) * e;
  dN[16 + 1] = -(z + 1.0) * (t + 1.0) * e;
  dN[16 + 2] = (z + 1.0) * (t + 1.0) * e;
  dN[16 + 3] = -(z - 1.0) * (t + 1.0) * e;
  dN[16 + 4] = -(z - 1.0) * (t - 1.0) * e;
  dN[16 + 5] = (z + 1.0) * (t - 1.0) * e;
  dN[16 + 6] = -(z + 1.0) * (t - 1.0) * e;
  dN[16 + 7] = (z - 1.0) * (t - 1.0) * e;
}

// local coordinates of p in the hex with corners X[a*3+k]; true when inside
// the widened reference cube
__device__ bool hex_local(const double *X, const double p[3], double xi[3]) {
#pragma clang fp contract(off)
  double N[8], dN[24], J[9], Ji[9], r[3];
  xi[0] = xi[1] = xi[2] = 0.0;
  for (int it = 0; it < kHexNewton; ++it) {
    hex_shape(xi, N, dN);
    for (int k = 0; k < 3; ++k) {
      double x = 0.0;
      for (int a = 0; a < 8; ++a) x = x + N[a] * X[a * 3 + k];
      r[k] = x - p[k];
      for (int d = 0; d < 3; ++d) {
        double j = 0.0;
        for (int a = 0; a < 8; ++a) j = j + dN[a + 8 * d] * X[a * 3 + k];
        OFD_M3(J, k, d) = j;
      }
    }
    inverse3x3(J, Ji);
    for (int d = 0; d < 3; ++d)
      xi[d] = xi[d] - (OFD_M3(Ji, d, 0) * r[0] + OFD_M3(Ji, d, 1) * r[1] +
                       OFD_M3(Ji, d, 2) * r[2]);
  }
  const double lim = 1.0 + kHexTol;
  return fabs(xi[0]) <= lim && fabs(xi[1]) <= lim && fabs(xi[2]) <= lim;
}

}  // namespace

// Per element: the widened bounding box (6), whether it is a source element,
// and its geometry: tets the inverse edge matrix of interpolation.f90:36-43
// with the first corner's x and d, hexes the corners x + d.
template <int NSHL>
__global__ void __launch_bounds__(256)
k_fringe_prep(int E, int N, const int32_t *__restrict__ ien, const int32_t *__restrict__ elm_id,
              int32_t src_id, const double *__restrict__ xg, const double *__restrict__ dg,
              double *__restrict__ box, int8_t *__restrict__ ok, double *__restrict__ geo) {
#pragma clang fp contract(off)
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const bool src = elm_id[e] == src_id;
  ok[e] = src ? 1 : 0;
  if (!src) return;
  double xl[NSHL][3], dl[NSHL][3], X[NSHL][3];
  for (int a = 0; a < NSHL; ++a) {
    const int n = ien[(size_t)a * E + e];
    for (int k = 0; k < 3; ++k) {
      xl[a][k] = xg[(size_t)k * N + n];
      dl[a][k] = dg[(size_t)k * N + n];
      X[a][k] = xl[a][k] + dl[a][k];
    }
  }
  double lo[3], hi[3], ext = 0.0;
  for (int k = 0; k < 3; ++k) {
    lo[k] = X[0][k];
    hi[k] = X[0][k];
    for (int a = 1; a < NSHL; ++a) {
      lo[k] = X[a][k] < lo[k] ? X[a][k] : lo[k];
      hi[k] = X[a][k] > hi[k] ? X[a][k] : hi[k];
    }
    const double w = hi[k] - lo[k];
    ext = w > ext ? w : ext;
  }
  const double m = kBoxMargin * ext;
  for (int k = 0; k < 3; ++k) {
    box[(size_t)e * 6 + k] = lo[k] - m;
    box[(size_t)e * 6 + 3 + k] = hi[k] + m;
  }
  if (NSHL == 4) {
    double A[9], Ai[9];
    for (int k = 0; k < 3; ++k)
      for (int a = 1; a < 4; ++a) OFD_M3(A, k, a - 1) = xl[a][k] - xl[0][k] + dl[a][k] - dl[0][k];
    inverse3x3(A, Ai);
    double *g = geo + (size_t)e * kTetGeo;
    for (int t = 0; t < 9; ++t) g[t] = Ai[t];
    for (int k = 0; k < 3; ++k) {
      g[9 + k] = xl[0][k];
      g[12 + k] = dl[0][k];
    }
  } else {
    double *g = geo + (size_t)e * kHexGeo;
    for (int a = 0; a < 8; ++a)
      for (int k = 0; k < 3; ++k) g[a * 3 + k] = X[a][k];
  }
}

// One wavefront per receiver j: donor[j] = first source element holding the
// point (-1 if none), val(nrecv,3) column-major = interpolated ug.
template <int NSHL>
__global__ void __launch_bounds__(256)
k_fringe_search(int E, int N, int nrecv, const double *__restrict__ xr,
                const double *__restrict__ dr, const double *__restrict__ box,
                const int8_t *__restrict__ ok, const double *__restrict__ geo,
                const int32_t *__restrict__ ien, const double *__restrict__ ug,
                int32_t *__restrict__ donor, double *__restrict__ val) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (j >= nrecv) return;  // uniform over the wave
  double x[3], d[3], p[3];
  for (int k = 0; k < 3; ++k) {
    x[k] = xr[(size_t)k * nrecv + j];
    d[k] = dr[(size_t)k * nrecv + j];
    p[k] = x[k] + d[k];
  }
  for (int base = 0; base < E; base += 64) {
    const int e = base + lane;
    bool hit = false;
    double w[4] = {0.0, 0.0, 0.0, 0.0};
    if (e < E && ok[e]) {
      const double *b = box + (size_t)e * 6;
      if (NSHL == 4) {
        // the point the reference's test places: x_1 + d_1 + bvec
        const double *g = geo + (size_t)e * kTetGeo;
        const double q0 = p[0] + g[12], q1 = p[1] + g[13], q2 = p[2] + g[14];
        if (q0 >= b[0] && q0 <= b[3] && q1 >= b[1] && q1 <= b[4] && q2 >= b[2] && q2 <= b[5]) {
          double bv[3];
          for (int k = 0; k < 3; ++k) bv[k] = x[k] - g[9 + k] + d[k];
          for (int i = 0; i < 3; ++i)
            w[i] = OFD_M3(g, i, 0) * bv[0] + OFD_M3(g, i, 1) * bv[1] + OFD_M3(g, i, 2) * bv[2];
          w[3] = 1.0 - (w[0] + w[1] + w[2]);
          hit = w[0] >= 0 && w[1] >= 0 && w[2] >= 0 && w[3] >= 0;
        }
      } else {
        if (p[0] >= b[0] && p[0] <= b[3] && p[1] >= b[1] && p[1] <= b[4] && p[2] >= b[2] &&
            p[2] <= b[5])
          hit = hex_local(geo + (size_t)e * kHexGeo, p, w);
      }
    }
    const unsigned long long m = __ballot(hit);
    if (m) {
      const int first = __ffsll(m) - 1;
      if (lane == first) {
        double Nw[8];
        if (NSHL == 4) {
          for (int a = 0; a < 4; ++a) Nw[a] = w[a];
        } else {
          double dN[24];
          hex_shape(w, Nw, dN);
        }
        for (int k = 0; k < 3; ++k) {
          double s = 0.0;
          for (int a = 0; a < NSHL; ++a)
            s = s + Nw[a] * ug[(size_t)k * N + ien[(size_t)a * E + e]];
          val[(size_t)k * nrecv + j] = s;
        }
        donor[j] = e;
      }
      return;
    }
  }
  if (lane == 0) {
    donor[j] = -1;
    for (int k = 0; k < 3; ++k) val[(size_t)k * nrecv + j] = 0.0;
  }
}

// setBCs_overset step 3 (setBCs.f90:158-165): node k = nodes[j] becomes
// Dirichlet in u, v, w with ug = val(j) and acg consistent with it
__global__ void k_fringe_apply(int nrecv, int N, const int32_t *__restrict__ nodes,
                               const double *__restrict__ val, double dtgl, double gami,
                               uint8_t *__restrict__ ibc, double *__restrict__ ug,
                               const double *__restrict__ ugold, double *__restrict__ acg,
                               const double *__restrict__ acgold) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nrecv) return;
  const int k = nodes[j];
  ibc[k] = (uint8_t)(ibc[k] | 7u);
  for (int c = 0; c < 3; ++c) {
    const double v = val[(size_t)c * nrecv + j];
    const size_t o = (size_t)c * N + k;
    ug[o] = v;
    acg[o] = ((v - ugold[o]) * dtgl + (gami - 1.0) * acgold[o]) / gami;
  }
}

#undef OFD_M3

// ---- the donor search without allocation (ofd_fringe_interp_noalloc) ----
//
// The search's scratch -- per element the widened box, the source flag and
// the geometry, per receiver the staged coordinates, donors and values --
// lives in one file-scope device array of kFringeScratchBytes, one copy per
// GPU for the life of the process.  Its device address is looked up once per
// device at context creation (hipGetSymbolAddress) and used from then on;
// host code never passes the array's own name to a kernel.  Every region is
// written in full by the kernels before it is read (prep writes the flag of
// every element, the search writes every receiver's donor and value), so
// nothing is cleared between calls.  Calls on one GPU share the array: calls
// that overlap in time on one device (two contexts on two streams) race.

namespace {
constexpr int kMaxScratchDevices = 64;
__device__ __align__(256) unsigned char scratch[kFringeScratchBytes];
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxScratchDevices] = {nullptr};

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
}  // namespace

int fringe_scratch_setup(int device) {
  OFD_CHECK(device >= 0 && device < kMaxScratchDevices, OFD_ERR_UNSUPPORTED,
            "device ordinal beyond the fringe scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return OFD_OK;
  void *ptr = nullptr;
  OFD_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  OFD_CHECK(ptr != nullptr, OFD_ERR_HIP, "hipGetSymbolAddress gave no address for the scratch");
  g_scratch_ptr[device] = ptr;
  return OFD_OK;
}

size_t fringe_scratch_need(int nshl, int nelem, int nrecv) {
  const int geo = nshl == 4 ? kTetGeo : kHexGeo;
  return align256((size_t)nelem * 6 * sizeof(double)) +
         align256((size_t)nelem * geo * sizeof(double)) + align256((size_t)nelem) +
         3 * align256((size_t)nrecv * 3 * sizeof(double)) +
         align256((size_t)nrecv * sizeof(int32_t));
}

}  // namespace ofd

using namespace ofd;

extern "C" int ofd_fringe_interp_noalloc(ofd_ctx *x, int32_t src_id, int nrecv,
                                         const double *xr, const double *dr, int32_t *donor,
                                         double *val) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  OFD_CHECK(nrecv >= 0, OFD_ERR_ARG, "nrecv < 0");
  Ctx &c = x->c;
  OFD_CHECK(c.nnode > 0 && c.d_ien, OFD_ERR_ARG, "no mesh uploaded");
  if (nrecv == 0) return OFD_OK;
  OFD_CHECK(xr && dr && donor && val, OFD_ERR_ARG, "null argument");
  OFD_CHECK(c.nshl == 4 || c.nshl == 8, OFD_ERR_UNSUPPORTED,
            "the donor search is implemented for linear tets and trilinear hexes");
  const size_t need = fringe_scratch_need(c.nshl, c.nelem, nrecv);
  OFD_CHECK(need <= kFringeScratchBytes, OFD_ERR_ARG,
            "fringe search needs " + std::to_string(need) + " bytes of scratch for " +
                std::to_string(c.nelem) + " elements and " + std::to_string(nrecv) +
                " receivers; the static scratch holds " + std::to_string(kFringeScratchBytes));
  void *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (c.device >= 0 && c.device < kMaxScratchDevices) base = g_scratch_ptr[c.device];
  }
  OFD_CHECK(base, OFD_ERR_ARG, "fringe scratch not set up for this device");
  OFD_HIP(hipSetDevice(c.device));
  const int geo_w = c.nshl == 4 ? kTetGeo : kHexGeo;
  unsigned char *p = static_cast<unsigned char *>(base);
  double *box = reinterpret_cast<double *>(p);
  p += align256((size_t)c.nelem * 6 * sizeof(double));
  double *geo = reinterpret_cast<double *>(p);
  p += align256((size_t)c.nelem * geo_w * sizeof(double));
  int8_t *ok = reinterpret_cast<int8_t *>(p);
  p += align256((size_t)c.nelem);
  const size_t vb = (size_t)nrecv * 3 * sizeof(double);
  double *sxr = reinterpret_cast<double *>(p);
  p += align256(vb);
  double *sdr = reinterpret_cast<double *>(p);
  p += align256(vb);
  double *sval = reinterpret_cast<double *>(p);
  p += align256(vb);
  int32_t *sdonor = reinterpret_cast<int32_t *>(p);
  // pageable host memory on both sides: no staging buffer of the library's
  OFD_HIP(hipMemcpyAsync(sxr, xr, vb, hipMemcpyHostToDevice, c.stream));
  OFD_HIP(hipMemcpyAsync(sdr, dr, vb, hipMemcpyHostToDevice, c.stream));
  OFD_HIP(launch_fringe(c, src_id, nrecv, sxr, sdr, sdonor, sval, box, ok, geo));
  OFD_HIP(hipMemcpyAsync(donor, sdonor, (size_t)nrecv * sizeof(int32_t), hipMemcpyDeviceToHost,
                         c.stream));
  OFD_HIP(hipMemcpyAsync(val, sval, vb, hipMemcpyDeviceToHost, c.stream));
  OFD_HIP(hipStreamSynchronize(c.stream));
  for (int j = 0; j < nrecv; ++j) donor[j] += 1;  // 1-based, 0 = no donor
  return OFD_OK;
}

namespace ofd {

hipError_t launch_fringe(Ctx &c, int32_t src_id, int nrecv, const double *xr,
                         const double *dr, int32_t *donor, double *val, double *box,
                         int8_t *ok, double *geo) {
  const int E = c.nelem, N = c.nnode;
  const dim3 gp((E + 255) / 256), gs((nrecv + 3) / 4);
  const double *xg = c.d_field[OFD_XG], *dg = c.d_field[OFD_DG], *ug = c.d_field[OFD_UG];
  if (c.nshl == 4) {
    hipLaunchKernelGGL(k_fringe_prep<4>, gp, dim3(256), 0, c.stream, E, N, c.d_ien,
                       c.d_elm_id, src_id, xg, dg, box, ok, geo);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fringe_search<4>, gs, dim3(256), 0, c.stream, E, N, nrecv, xr, dr,
                       box, ok, geo, c.d_ien, ug, donor, val);
  } else {
    hipLaunchKernelGGL(k_fringe_prep<8>, gp, dim3(256), 0, c.stream, E, N, c.d_ien,
                       c.d_elm_id, src_id, xg, dg, box, ok, geo);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fringe_search<8>, gs, dim3(256), 0, c.stream, E, N, nrecv, xr, dr,
                       box, ok, geo, c.d_ien, ug, donor, val);
  }
  return hipGetLastError();
}

hipError_t launch_fringe_apply(Ctx &c, int nrecv, const int32_t *nodes, const double *val) {
  double **f = c.d_field;
  hipLaunchKernelGGL(k_fringe_apply, dim3((nrecv + 255) / 256), dim3(256), 0, c.stream, nrecv,
                     c.nnode, nodes, val, 1.0 / c.prm[OFD_P_DELT], c.prm[OFD_P_GAMI], c.d_ibc,
                     f[OFD_UG], f[OFD_UGOLD], f[OFD_ACG], f[OFD_ACGOLD]);
  return hipGetLastError();
}

}  // namespace ofd
