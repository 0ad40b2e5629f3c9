// C ABI of the gfx950 OversetDevFort hot-path library (declared in
// include/ofd.h).  Host-side bookkeeping: context and device buffers, the
// sparsity pattern (genSparStruc.f90:4-105), element colouring, the
// (element, a, b) -> nonzero scatter map that replaces the binary search of
// SparseMatLoc_3D (FillSparseMat.f90:4-28), and the GMRES / Newton drivers
// that sequence the kernels.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <new>

#include "ofd_internal.h"

namespace ofd {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

template <typename T>
static int dalloc(Ctx &c, T **p, size_t n) {
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  if (n == 0) return OFD_OK;
  hipError_t e = hipMalloc(reinterpret_cast<void **>(p), n * sizeof(T));
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
    return OFD_ERR_ALLOC;
  }
  c.dev_bytes += (int64_t)(n * sizeof(T));
  return OFD_OK;
}

template <typename T>
static void dfree(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

static int ensure_pin(Ctx &c, size_t bytes) {
  if (c.h_pin_bytes >= bytes) return OFD_OK;
  if (c.h_pin) (void)hipHostFree(c.h_pin);
  c.h_pin = nullptr;
  c.h_pin_bytes = 0;
  OFD_HIP(hipHostMalloc(reinterpret_cast<void **>(&c.h_pin), bytes));
  c.h_pin_bytes = bytes;
  return OFD_OK;
}

// host -> device through the pinned staging buffer
static int h2d(Ctx &c, void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return OFD_OK;
  int rc = ensure_pin(c, bytes);
  if (rc) return rc;
  std::memcpy(c.h_pin, src, bytes);
  OFD_HIP(hipMemcpyAsync(dst, c.h_pin, bytes, hipMemcpyHostToDevice, c.stream));
  OFD_HIP(hipStreamSynchronize(c.stream));
  return OFD_OK;
}

static int d2h(Ctx &c, void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return OFD_OK;
  int rc = ensure_pin(c, bytes);
  if (rc) return rc;
  OFD_HIP(hipMemcpyAsync(c.h_pin, src, bytes, hipMemcpyDeviceToHost, c.stream));
  OFD_HIP(hipStreamSynchronize(c.stream));
  std::memcpy(dst, c.h_pin, bytes);
  return OFD_OK;
}

static int field_len(const Ctx &c, int f) {
  const bool vec3 = f <= OFD_ACGMOLD || f == OFD_SOLU || f == OFD_LAMU;
  return vec3 ? 3 * c.nnode : c.nnode;
}

static int need_mesh(const Ctx &c) {
  OFD_CHECK(c.nnode > 0 && c.nelem > 0, OFD_ERR_ARG, "no mesh uploaded");
  return OFD_OK;
}

static int need_pattern(const Ctx &c) {
  OFD_CHECK(c.icnt > 0 && c.d_col, OFD_ERR_ARG, "no sparsity pattern (ofd_build_pattern)");
  return OFD_OK;
}

// LHSTem transposed and its solver copies (sized by the pattern)
static void free_tem_transpose(Ctx &c) {
  dfree(c.tT_A);
  dfree(c.tT_A32);
  dfree(c.tT_P);
  dfree(c.tT_Ap);
  dfree(c.tT_Ap32);
  c.tT_mat_valid = c.tT_prec_valid = c.tT_a32_valid = c.tT_ap32_valid = false;
}

// NS + VOF adjoint: the transposed matrix and its solver copies, vectors
static void free_nsvof_adjoint(Ctx &c) {
  dfree(c.t_AV);
  dfree(c.t_AV32);
  dfree(c.t_PV);
  dfree(c.t_ApV);
  dfree(c.t_ApV32);
  c.t_matV_valid = c.t_precV_valid = c.t_pV_valid = c.t_aV32_valid = c.t_apV32_valid = false;
  c.t_nsv_icnt = 0;
  dfree(c.d_lamV);
  dfree(c.d_lamVb);
  dfree(c.d_lamVd);
  dfree(c.d_jtV);
  dfree(c.d_vjpV);
  c.vadj_n4 = 0;
}

// temperature adjoint vectors (sized by the mesh)
static void free_tem_adjoint(Ctx &c) {
  dfree(c.d_lamT);
  dfree(c.d_lamTb);
  dfree(c.d_lamTd);
  dfree(c.d_jtT);
  dfree(c.d_vjpT);
  c.tadj_n4 = 0;
}

// the element assemblies' per-element point data (1536 bytes per hex, 1088
// per prism, 768 per tet), allocated at the first assembly
static int ensure_pts(Ctx &c) {
  if (c.nshl != 8 && c.nshl != 4 && c.nshl != 6) return OFD_OK;
  const size_t len = (size_t)c.nelem * pts_stride(c.nshl);
  if (c.d_pts && c.pts_len == len) return OFD_OK;
  int rc = dalloc(c, &c.d_pts, len);
  c.pts_len = rc ? 0 : len;
  return rc;
}

// Krylov workspace for restart length K
static int ensure_krylov(Ctx &c, int K) {
  OFD_CHECK(K >= 1 && K <= kMaxKspace, OFD_ERR_ARG, "kspace out of range [1, 64]");
  if (!c.h_flags) {
    // coherent: the device's flag stores reach the host while kernels run
    OFD_HIP(hipHostMalloc(reinterpret_cast<void **>(&c.h_flags),
                          sizeof(int32_t) * 2 * kMaxKspace, hipHostMallocCoherent));
    OFD_HIP(hipHostMalloc(reinterpret_cast<void **>(&c.h_rec), sizeof(CycleRecord) * 2,
                          hipHostMallocCoherent));
    std::memset(c.h_rec, 0, sizeof(CycleRecord) * 2);
  }
  // vectors of the largest system seen so far (NS, T or NSVOF tuples)
  const int nt = c.gm_n > c.kalloc_n ? c.gm_n : c.kalloc_n;
  if (nt > c.kalloc_n) {
    dfree(c.d_W);
    dfree(c.d_W32);
    c.kalloc = c.kalloc32 = 0;
    c.kalloc_n = nt;
  }
  if (c.basis32) {
    if (c.kalloc32 >= K) return OFD_OK;
    int rc = dalloc(c, &c.d_W32, (size_t)(K + 1) * c.kalloc_n * 4);
    if (rc) return rc;
    c.kalloc32 = K;
    return OFD_OK;
  }
  if (c.kalloc >= K) return OFD_OK;
  int rc = dalloc(c, &c.d_W, (size_t)(K + 1) * c.kalloc_n * 4);
  if (rc) return rc;
  c.kalloc = K;
  return OFD_OK;
}

// FP32 copy of the system matrix (plain SpMV in FP32)
static int ensure_a32(Ctx &c) {
  OFD_CHECK(c.mat_valid, OFD_ERR_ARG, "no assembled matrix");
  if (c.a32_valid) return OFD_OK;
  if (!c.d_A32) {
    int rc = dalloc(c, &c.d_A32, (size_t)c.icnt * 16);
    if (rc) return rc;
  }
  OFD_HIP(launch_convert_a32(c));
  c.a32_valid = true;
  return OFD_OK;
}

// block-Jacobi inverse and the left-preconditioned matrix P^-1 A (FP64, and
// FP32 when the mixed solver asks for it) of the current matrix; for the
// right form the inverse and the plain matrix (FP32 copy for the mixed
// solver)
static int ensure_prec(Ctx &c, int precision, bool right = false) {
  OFD_CHECK(c.mat_valid, OFD_ERR_ARG, "no assembled matrix");
  const bool want32 = precision == OFD_FP32;
  if (right) {
    int rc = ensure_prec(c, OFD_FP64);
    if (rc) return rc;
    return want32 ? ensure_a32(c) : OFD_OK;
  }
  if (c.prec_valid && (!want32 || c.ap32_valid)) return OFD_OK;
  if (!c.d_Ap64) {
    int rc = dalloc(c, &c.d_Ap64, (size_t)c.icnt * 16);
    if (rc) return rc;
  }
  if (want32 && !c.d_Ap32) {
    int rc = dalloc(c, &c.d_Ap32, (size_t)c.icnt * 16);
    if (rc) return rc;
  }
  OFD_HIP(launch_build_prec(c));
  OFD_HIP(launch_precondition_matrix(c, want32));
  c.prec_valid = true;
  c.ap32_valid = want32;
  return OFD_OK;
}

// read back the sampled Arnoldi-product timings recorded since the last
// host synchronisation (the caller has synchronised the stream)
constexpr int kProbeSlots = 4096;

static int probe_harvest(Ctx &c) {
  for (int i = 0; i < c.probe_used; ++i) {
    float t = 0.f;
    OFD_HIP(hipEventElapsedTime(&t, c.probe_ev[2 * i], c.probe_ev[2 * i + 1]));
    c.probe_ms += t;
    ++c.probe_count;
  }
  c.probe_used = 0;
  return OFD_OK;
}

// Wait until the device has written the convergence flag of Arnoldi step
// `idx` (-1 = not yet).  The kernel that closes step idx writes it, also
// when it returns at once behind a stopped cycle.  The stream is queried
// only every 2^20 spins (tens of ms), as a guard: a query enqueues a marker
// whose cache flush stalls the queue for several microseconds, and polling
// it every few hundred spins put such a bubble into about half of the
// Arnoldi steps.  An idle stream with the flag unwritten is reported as
// converged.
static int wait_flag(Ctx &c, int idx, bool *done) {
  volatile int32_t *f = c.cur_flags;
  for (unsigned spin = 1;; ++spin) {
    const int32_t v = f[idx];
    if (v >= 0) {
      *done = v != 0;
      return OFD_OK;
    }
    if ((spin & ((1u << 20) - 1u)) == 0u) {
      const hipError_t e = hipStreamQuery(c.stream);
      if (e == hipSuccess) {
        *done = f[idx] != 0;
        return OFD_OK;
      }
      if (e != hipErrorNotReady) OFD_HIP(e);
    }
    __builtin_ia32_pause();
  }
}

// Wait for the record of the cycle numbered `seq` in pinned slot `slot`
// (written by the outer residual's norm kernel) and copy it out.  No event
// and no copy on the stream: a device-to-host copy between two cycles held
// the queue for about 20 us.  The stream is queried only as a guard
// against a record that can no longer come.
static int wait_record(Ctx &c, int slot, int32_t seq, CycleRecord *out) {
  volatile CycleRecord *r = c.h_rec + slot;
  for (unsigned spin = 1;; ++spin) {
    if (r->seq == seq) {
      std::atomic_thread_fence(std::memory_order_acquire);
      out->resid = r->resid;
      out->rnorm0 = r->rnorm0;
      out->rnorm = r->rnorm;
      out->iters = r->iters;
      out->seq = seq;
      return OFD_OK;
    }
    if ((spin & ((1u << 20) - 1u)) == 0u) {
      const hipError_t e = hipStreamQuery(c.stream);
      if (e == hipSuccess && r->seq != seq) {
        set_error("solve: cycle record not written");
        return OFD_ERR_HIP;
      }
      if (e != hipSuccess && e != hipErrorNotReady) OFD_HIP(e);
    }
    __builtin_ia32_pause();
  }
}

// the T system's preconditioner: inverse diagonal and D^-1 LHSTem (FP64,
// plus FP32 for the mixed solver)
static int ensure_prec_s(Ctx &c, int precision) {
  OFD_CHECK(c.matT_valid, OFD_ERR_ARG, "no assembled temperature matrix");
  const bool want32 = precision == OFD_FP32;
  if (c.precT_valid && (!want32 || c.apT32_valid)) return OFD_OK;
  if (!c.d_ApT) {
    int rc = dalloc(c, &c.d_ApT, (size_t)c.icnt);
    if (rc) return rc;
  }
  if (want32 && !c.d_ApT32) {
    int rc = dalloc(c, &c.d_ApT32, (size_t)c.icnt);
    if (rc) return rc;
  }
  OFD_HIP(launch_build_prec_s(c, want32));
  c.precT_valid = true;
  c.apT32_valid = want32;
  return OFD_OK;
}

// Points the GMRES kernels at the T system (ceil(N/4) 4-tuples) for the
// lifetime of the guard; the NS node vectors otherwise.
struct TemSystem {
  Ctx &c;
  explicit TemSystem(Ctx &cc) : c(cc) {
    c.gm_kind = 1;
    c.gm_n = (c.nnode + 3) / 4;
  }
  ~TemSystem() {
    c.gm_kind = 0;
    c.gm_n = c.nnode;
  }
};

// the NSVOF system's vectors ([N][4] then [N] phi) as ceil(5N/4) 4-tuples
struct NsvofSystem {
  Ctx &c;
  explicit NsvofSystem(Ctx &cc) : c(cc) {
    c.gm_kind = 2;
    c.gm_n = (5 * c.nnode + 3) / 4;
  }
  ~NsvofSystem() {
    c.gm_kind = 0;
    c.gm_n = c.nnode;
  }
};

// FP32 copy of the NSVOF matrix
static int ensure_aV32(Ctx &c) {
  OFD_CHECK(c.matV_valid, OFD_ERR_ARG, "no assembled NSVOF matrix");
  if (c.aV32_valid) return OFD_OK;
  if (!c.d_AV32) {
    int rc = dalloc(c, &c.d_AV32, (size_t)c.icnt * 24);
    if (rc) return rc;
  }
  OFD_HIP(launch_convert_aV32(c));
  c.aV32_valid = true;
  return OFD_OK;
}

// the NSVOF preconditioner: block inverses and P^-1 A (FP64, FP32); for the
// right form the block inverses and the plain matrix (FP32 copy for the
// mixed solver)
static int ensure_prec_v(Ctx &c, int precision, bool right) {
  OFD_CHECK(c.matV_valid, OFD_ERR_ARG, "no assembled NSVOF matrix");
  const bool want32 = precision == OFD_FP32;
  if (right) {
    if (!c.pV_valid) {
      OFD_HIP(launch_build_prec_v(c, false, false));
      c.pV_valid = true;
    }
    return want32 ? ensure_aV32(c) : OFD_OK;
  }
  if (c.precV_valid && (!want32 || c.apV32_valid)) return OFD_OK;
  if (!c.d_ApV) {
    int rc = dalloc(c, &c.d_ApV, (size_t)c.icnt * 24);
    if (rc) return rc;
  }
  if (want32 && !c.d_ApV32) {
    int rc = dalloc(c, &c.d_ApV32, (size_t)c.icnt * 24);
    if (rc) return rc;
  }
  OFD_HIP(launch_build_prec_v(c, want32, true));
  c.precV_valid = c.pV_valid = true;
  c.apV32_valid = want32;
  return OFD_OK;
}

// one GMRES cycle on right-hand side b (device [N][4]); x += increment.
// flag_set selects the step-flag set (cycles queued back to back alternate).
static int gmres_cycle(Ctx &c, int precision, int ortho, const double *b, double tol,
                       int K, int kmin, double *x, int *iters, double *resid,
                       double *rnorm0, bool sync = true, int flag_set = 0) {
  const bool right = (ortho & OFD_PREC_RIGHT) != 0;
  ortho &= ~OFD_PREC_RIGHT;
  OFD_CHECK(ortho == OFD_ORTHO_MGS || ortho == OFD_ORTHO_CGS2 || ortho == OFD_ORTHO_CGS1,
            OFD_ERR_ARG, "bad ortho");
  OFD_CHECK(!right || (c.gm_kind != 1 && ortho != OFD_ORTHO_MGS), OFD_ERR_ARG,
            "OFD_PREC_RIGHT applies to the NS and NSVOF systems with OFD_ORTHO_CGS1 or CGS2");
  c.right = right;
  if (right && c.t_n < c.gm_n) {
    // T = P^-1 W_k, the right form's Arnoldi product input
    const size_t np = (size_t)c.gm_n * 4;
    int rc = dalloc(c, &c.d_T, np);
    if (!rc) rc = dalloc(c, &c.d_T32, np);
    c.t_n = rc ? 0 : c.gm_n;
    if (rc) return rc;
  }
  // the mixed CGS solver keeps its Krylov basis in FP32
  c.basis32 = precision == OFD_FP32 && ortho != OFD_ORTHO_MGS;
  int rc = ensure_krylov(c, K);
  if (rc) return rc;
  rc = c.gm_kind == 1 ? ensure_prec_s(c, precision)
                      : (c.gm_kind == 2 ? ensure_prec_v(c, precision, right)
                                        : ensure_prec(c, precision, right));
  if (rc) return rc;
  OFD_HIP(launch_gmres_init(c, b, tol, K, kmin));
  // The host enqueues iterations ahead of the device and learns about
  // convergence through a flag the device writes to coherent pinned host
  // memory when it closes each step; it waits only on the flag of step
  // k - kLag - 1, so the device queue never drains, and puts no marker
  // between the kernels of an iteration.  Kernels launched after
  // convergence return at once.
  constexpr int kLag = 3;
  c.cur_flags = c.h_flags + (size_t)(flag_set & 1) * kMaxKspace;
  for (int k = 0; k < K; ++k) c.cur_flags[k] = -1;
  int klast = -1;
  for (int k = 0; k < K; ++k) {
    // the first orthogonalisation kernel of iteration j closes step j-1
    if (k >= kLag + 1) {
      bool stop = false;
      rc = wait_flag(c, k - kLag - 1, &stop);
      if (rc) return rc;
      if (stop) break;
    }
    // sampled by a running launch count, so every step index k is covered
    const bool probe = c.probe_every > 0 && (c.probe_tick++ % c.probe_every) == 0 &&
                       c.probe_used < (int)c.probe_ev.size() / 2;
    if (probe) {
      OFD_HIP(launch_spmv_arnoldi(c, precision, k, c.probe_ev[2 * c.probe_used],
                                  c.probe_ev[2 * c.probe_used + 1]));
      ++c.probe_used;
    } else {
      OFD_HIP(launch_spmv_arnoldi(c, precision, k));
    }
    if (ortho == OFD_ORTHO_MGS) {
      for (int j = 0; j <= k + 1; ++j) OFD_HIP(launch_mgs_step(c, k, j));
    } else {
      // (product above); projection dots (closing step k-1); update sweep(s)
      OFD_HIP(launch_cgs_dots(c, k));
      if (ortho == OFD_ORTHO_CGS1) {
        OFD_HIP(launch_cgs_update(c, k, 1));
      } else {
        OFD_HIP(launch_cgs_update(c, k, 2));
        OFD_HIP(launch_cgs_update(c, k, 3));
      }
    }
    klast = k;
  }
  if (klast >= 0) OFD_HIP(launch_gmres_close(c, klast, ortho));
  OFD_HIP(launch_gmres_finish(c, x));
  if (!sync) return OFD_OK;  // the caller reads the cycle scalars later
  OFD_HIP(hipMemcpyAsync(c.h_gs, c.d_gs, sizeof(GmresState), hipMemcpyDeviceToHost,
                         c.stream));
  OFD_HIP(hipStreamSynchronize(c.stream));
  rc = probe_harvest(c);
  if (rc) return rc;
  if (iters) *iters = c.h_gs->iters;
  if (resid) *resid = c.h_gs->resid;
  if (rnorm0) *rnorm0 = c.h_gs->rnorm0;
  if (!std::isfinite(c.h_gs->resid) || !std::isfinite(c.h_gs->rnorm0)) {
    set_error("GMRES: non-finite residual");
    return OFD_ERR_NONFINITE;
  }
  return OFD_OK;
}

static int norm2_dev(Ctx &c, const double *v, double *out) {
  OFD_HIP(launch_norm2(c, v, 0));
  OFD_HIP(hipMemcpyAsync(c.h_gs, c.d_gs, sizeof(GmresState), hipMemcpyDeviceToHost,
                         c.stream));
  OFD_HIP(hipStreamSynchronize(c.stream));
  *out = c.h_gs->scal[0];
  return OFD_OK;
}

}  // namespace ofd

using namespace ofd;

extern "C" {

const char *ofd_last_error(void) { return g_err.c_str(); }

int ofd_version(void) { return 1; }

int ofd_ctx_create(int device, ofd_ctx **out) {
  OFD_CHECK(out, OFD_ERR_ARG, "null output");
  int n = 0;
  OFD_HIP(hipGetDeviceCount(&n));
  OFD_CHECK(device >= 0 && device < n, OFD_ERR_ARG, "device ordinal out of range");
  OFD_HIP(hipSetDevice(device));
  ofd_ctx *x = new (std::nothrow) ofd_ctx();
  OFD_CHECK(x, OFD_ERR_ALLOC, "out of host memory");
  Ctx &c = x->c;
  c.device = device;
  if (hipDeviceGetAttribute(&c.n_cu, hipDeviceAttributeMultiprocessorCount, device) !=
          hipSuccess ||
      c.n_cu < 1)
    c.n_cu = 256;
  hipError_t e = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete x;
    set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
    return OFD_ERR_HIP;
  }
  // the device address of the no-allocation fringe search's static scratch
  int rc = fringe_scratch_setup(device);
  if (rc) {
    ofd_ctx_destroy(x);
    return rc;
  }
  rc = dalloc(c, &c.d_gs, 1);
  if (!rc) rc = dalloc(c, &c.d_part, (size_t)4 * kMaxRedBlocks * (kMaxKspace + 1));
  if (!rc) rc = dalloc(c, &c.d_negdet, 1);
  if (rc) {
    ofd_ctx_destroy(x);
    return rc;
  }
  e = hipHostMalloc(reinterpret_cast<void **>(&c.h_gs), sizeof(GmresState));
  if (e != hipSuccess) {
    ofd_ctx_destroy(x);
    set_error("hipHostMalloc failed");
    return OFD_ERR_ALLOC;
  }
  std::memset(c.h_gs, 0, sizeof(GmresState));
  e = hipMemcpy(c.d_gs, c.h_gs, sizeof(GmresState), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    ofd_ctx_destroy(x);
    set_error("hipMemcpy failed");
    return OFD_ERR_HIP;
  }
  c.prm[OFD_P_ALMI] = c.prm[OFD_P_ALFI] = c.prm[OFD_P_GAMI] = c.prm[OFD_P_BETI] = 1.0;
  c.prm[OFD_P_DELT] = 1.0;
  c.prm[OFD_P_RHO] = 1.0;
  c.prm[OFD_P_GZ] = -9.81;
  *out = x;
  return OFD_OK;
}

int ofd_ctx_destroy(ofd_ctx *x) {
  if (!x) return OFD_OK;
  Ctx &c = x->c;
  (void)hipSetDevice(c.device);
  if (c.stream) (void)hipStreamSynchronize(c.stream);
  dfree(c.d_ien);
  dfree(c.d_elm_id);
  dfree(c.d_col);
  dfree(c.d_row);
  dfree(c.d_emap);
  dfree(c.d_color_elems);
  dfree(c.d_first);
  dfree(c.d_inc_off);
  dfree(c.d_inc);
  dfree(c.d_pts);
  dfree(c.d_emap_c);
  dfree(c.d_first_c);
  dfree(c.d_cpos);
  dfree(c.d_cl_elem);
  dfree(c.d_cl_map);
  dfree(c.d_cl_kf);
  dfree(c.d_cl_nslot);
  dfree(c.d_fe_elem);
  dfree(c.d_fe_mask);
  dfree(c.d_diag);
  for (auto &f : c.d_field) dfree(f);
  dfree(c.d_ibc);
  dfree(c.d_bcval);
  dfree(c.d_noderec);
  dfree(c.d_A64);
  dfree(c.d_A32);
  dfree(c.d_P64);
  dfree(c.d_Ap64);
  dfree(c.d_Ap32);
  dfree(c.d_rhs);
  dfree(c.d_b);
  dfree(c.d_x);
  dfree(c.d_r);
  dfree(c.d_cfl);
  dfree(c.d_negdet);
  dfree(c.d_W);
  dfree(c.d_W32);
  dfree(c.d_part);
  dfree(c.d_gs);
  dfree(c.d_tmap);
  dfree(c.t_A64);
  dfree(c.t_A32);
  dfree(c.t_P64);
  dfree(c.t_Ap64);
  dfree(c.t_Ap32);
  dfree(c.d_jt);
  dfree(c.d_lam);
  dfree(c.d_lamr);
  dfree(c.d_lamb);
  dfree(c.d_lamd);
  dfree(c.d_vjp);
  dfree(c.d_fv);
  dfree(c.d_fz);
  dfree(c.d_fh);
  c.fg_m = 0;
  c.fg_n = 0;
  free_tem_transpose(c);
  free_tem_adjoint(c);
  free_nsvof_adjoint(c);
  dfree(c.d_AV);
  dfree(c.d_AV32);
  dfree(c.d_PV);
  dfree(c.d_ApV);
  dfree(c.d_ApV32);
  dfree(c.d_T);
  dfree(c.d_T32);
  c.t_n = 0;
  dfree(c.d_rhsV);
  dfree(c.d_xV);
  dfree(c.d_rV);
  dfree(c.d_yV);
  dfree(c.d_AT);
  dfree(c.d_AT32);
  dfree(c.d_PT);
  dfree(c.d_ApT);
  dfree(c.d_ApT32);
  dfree(c.d_rhsT);
  dfree(c.d_xT);
  dfree(c.d_rT);
  dfree(c.d_yT);
  dfree(c.d_fr_box);
  dfree(c.d_fr_ok);
  dfree(c.d_fr_geo);
  dfree(c.d_fr_xr);
  dfree(c.d_fr_dr);
  dfree(c.d_fr_val);
  dfree(c.d_fr_donor);
  dfree(c.d_fr_nodes);
  if (c.h_gs) (void)hipHostFree(c.h_gs);
  if (c.h_flags) (void)hipHostFree(c.h_flags);
  if (c.h_rec) (void)hipHostFree(c.h_rec);
  for (auto &ev : c.probe_ev) (void)hipEventDestroy(ev);
  if (c.h_pin) (void)hipHostFree(c.h_pin);
  if (c.stream) (void)hipStreamDestroy(c.stream);
  delete x;
  return OFD_OK;
}

int ofd_upload_mesh(ofd_ctx *x, int nnode, int nelem, int nshl, const int32_t *ien,
                    const int32_t *elm_id, const double *xg) {
  OFD_CHECK(x && ien && elm_id && xg, OFD_ERR_ARG, "null argument");
  OFD_CHECK(nnode > 0 && nelem > 0, OFD_ERR_ARG, "empty mesh");
  OFD_CHECK(nnode <= kMaxRedBlocks * kRedBlock, OFD_ERR_UNSUPPORTED,
            "more than 1,048,576 nodes per context");
  OFD_CHECK(nshl == 8 || nshl == 4 || nshl == 6, OFD_ERR_UNSUPPORTED,
            "elements must be trilinear hexes (NSHL=8), linear tets (NSHL=4) or linear "
            "prisms (NSHL=6)");
  Ctx &c = x->c;
  OFD_HIP(hipSetDevice(c.device));
  const size_t ne = (size_t)nelem * nshl;
  c.h_ien.resize(ne);
  for (size_t t = 0; t < ne; ++t) {
    const int v = ien[t] - 1;
    OFD_CHECK(v >= 0 && v < nnode, OFD_ERR_ARG, "IEN entry out of range [1, NNODE]");
    c.h_ien[t] = v;
  }
  c.h_elm_id.assign(elm_id, elm_id + nelem);
  c.nnode = nnode;
  c.nelem = nelem;
  c.nshl = nshl;
  c.gm_n = nnode;
  c.gm_kind = 0;
  c.matV_valid = c.precV_valid = c.pV_valid = c.aV32_valid = c.apV32_valid = false;
  c.nsvV_n4 = 0;
  c.matT_valid = c.precT_valid = c.aT32_valid = c.apT32_valid = false;
  c.temT_n4 = 0;
  c.icnt = 0;
  c.mat_valid = c.prec_valid = c.a32_valid = c.ap32_valid = false;
  ++c.mat_version;
  ++c.matT_version;
  c.mus_on = false;
  c.ks_on = false;
  // the weak faces name elements of the old mesh
  c.h_face_e.clear();
  c.h_face_or.clear();
  c.fe_color_off.clear();
  c.fe_dirty = false;
  dfree(c.d_fe_elem);
  dfree(c.d_fe_mask);
  free_tem_transpose(c);
  free_tem_adjoint(c);
  free_nsvof_adjoint(c);
  // adjoint buffers are sized by the mesh and the pattern
  dfree(c.d_jt);
  dfree(c.d_lam);
  dfree(c.d_lamr);
  dfree(c.d_lamb);
  dfree(c.d_lamd);
  dfree(c.d_vjp);
  dfree(c.d_fv);
  dfree(c.d_fz);
  dfree(c.d_fh);
  c.fg_m = 0;
  c.fg_n = 0;
  dfree(c.t_P64);
  int rc = dalloc(c, &c.d_ien, ne);
  if (!rc) rc = dalloc(c, &c.d_elm_id, (size_t)nelem);
  for (int f = 0; f < OFD_NFIELDS && !rc; ++f)
    rc = dalloc(c, &c.d_field[f], (size_t)field_len(c, f));
  if (!rc) rc = dalloc(c, &c.d_ibc, (size_t)nnode);
  if (!rc) rc = dalloc(c, &c.d_bcval, (size_t)nnode * 3);
  if (!rc) rc = dalloc(c, &c.d_noderec, (size_t)nnode * 16);
  if (!rc) rc = dalloc(c, &c.d_rhs, (size_t)nnode * 4);
  if (!rc) rc = dalloc(c, &c.d_b, (size_t)nnode * 4);
  if (!rc) rc = dalloc(c, &c.d_x, (size_t)nnode * 4);
  if (!rc) rc = dalloc(c, &c.d_r, (size_t)nnode * 4);
  if (!rc) rc = dalloc(c, &c.d_cfl, (size_t)nelem);
  if (!rc) rc = dalloc(c, &c.d_P64, (size_t)nnode * 16);
  if (rc) return rc;
  rc = h2d(c, c.d_ien, c.h_ien.data(), ne * sizeof(int32_t));
  if (!rc) rc = h2d(c, c.d_elm_id, elm_id, (size_t)nelem * sizeof(int32_t));
  if (!rc) rc = h2d(c, c.d_field[OFD_XG], xg, (size_t)nnode * 3 * sizeof(double));
  if (rc) return rc;
  for (int f = 0; f < OFD_NFIELDS; ++f)
    if (f != OFD_XG)
      OFD_HIP(hipMemsetAsync(c.d_field[f], 0, (size_t)field_len(c, f) * sizeof(double),
                             c.stream));
  OFD_HIP(hipMemsetAsync(c.d_ibc, 0, (size_t)nnode, c.stream));
  OFD_HIP(hipMemsetAsync(c.d_bcval, 0, (size_t)nnode * 3 * sizeof(double), c.stream));
  OFD_HIP(hipMemsetAsync(c.d_x, 0, (size_t)nnode * 4 * sizeof(double), c.stream));
  OFD_HIP(hipStreamSynchronize(c.stream));
  return OFD_OK;
}

// Clusters of the hex pair kernel.  Greedy over the nodes, for a threshold
// from kNsClElems down to 1: a node with at least that many unclustered
// elements takes up to kNsClElems of them as a cluster (on a structured hex
// mesh the first pass tiles it with the 2 x 2 x 2 stars of every other
// node), each element while the cluster's distinct nonzeros stay within
// kNsClSlots.  Clusters are coloured first-fit so that clusters of one
// colour share no node; the first touch of each nonzero follows that order.
// More than 64 colours leaves the mesh to the per-element kernel.
static int build_ns_clusters(Ctx &c, const std::vector<int32_t> &order) {
  const int E = c.nelem, N = c.nnode, S = 8;
  OFD_CHECK(c.icnt < (int64_t(1) << 31), OFD_ERR_UNSUPPORTED,
            "more than 2^31 nonzero blocks");
  std::vector<int32_t> tpos(E);
  for (int t = 0; t < E; ++t) tpos[order[t]] = t;
  std::vector<int32_t> off((size_t)N + 1, 0), inc((size_t)E * S);
  for (int e = 0; e < E; ++e)
    for (int a = 0; a < S; ++a) off[(size_t)c.h_ien[(size_t)a * E + e] + 1]++;
  for (int i = 0; i < N; ++i) off[i + 1] += off[i];
  {
    std::vector<int32_t> pos(off.begin(), off.end() - 1);
    for (int e = 0; e < E; ++e)
      for (int a = 0; a < S; ++a) inc[pos[c.h_ien[(size_t)a * E + e]]++] = e;
  }
  auto node = [&](int e, int a) { return c.h_ien[(size_t)a * E + e]; };
  std::vector<uint8_t> done(E, 0);
  std::vector<int32_t> cl_elem;                  // [C][8] element ids, -1 none
  std::vector<uint16_t> cl_map;                  // [C][8][64]
  std::vector<int32_t> cl_nslot;
  std::vector<std::pair<int32_t, int32_t>> cl_pair;  // [C][kNsClSlots] (row, col) nodes
  int ln[kNsClElems * 8];
  uint16_t tab[kNsClElems * 8][kNsClElems * 8];
  std::vector<int32_t> cand;
  for (int thr = kNsClElems; thr >= 1; --thr) {
    for (int n = 0; n < N; ++n) {
      cand.clear();
      for (int k = off[n]; k < off[n + 1]; ++k)
        if (!done[inc[k]]) cand.push_back(inc[k]);
      if ((int)cand.size() < thr) continue;
      int nln = 0, nslot = 0, nm = 0;
      const size_t base = cl_elem.size();
      cl_elem.resize(base + kNsClElems, -1);
      cl_map.resize((base + kNsClElems) * 64, 0);
      cl_pair.resize((base / kNsClElems + 1) * kNsClSlots, {-1, -1});
      for (int e : cand) {
        if (nm == kNsClElems) break;
        // local node numbers (new ones appended) and the pairs it adds
        int li[8], nl = nln;
        for (int a = 0; a < S; ++a) {
          const int g = node(e, a);
          int j = 0;
          while (j < nl && ln[j] != g) ++j;
          if (j == nl) ln[nl++] = g;
          li[a] = j;
        }
        int add = 0;
        for (int b = 0; b < S; ++b)
          for (int a = 0; a < S; ++a)
            if (li[a] >= nln || li[b] >= nln || tab[li[a]][li[b]] == 0xffff) ++add;
        if (nslot + add > kNsClSlots) break;
        for (int j = nln; j < nl; ++j)
          for (int k = 0; k < nl; ++k) tab[j][k] = tab[k][j] = 0xffff;
        nln = nl;
        for (int b = 0; b < S; ++b)
          for (int a = 0; a < S; ++a) {
            uint16_t &sl = tab[li[a]][li[b]];
            if (sl == 0xffff) {
              sl = (uint16_t)nslot;
              cl_pair[base / kNsClElems * kNsClSlots + nslot] = {node(e, a), node(e, b)};
              ++nslot;
            }
            cl_map[(base + nm) * 64 + a + 8 * b] = sl;
          }
        cl_elem[base + nm] = e;
        done[e] = 1;
        ++nm;
      }
      cl_nslot.push_back(nslot);
    }
  }
  const int C = (int)cl_nslot.size();
  // colours: first fit over one 64-bit word per node
  std::vector<uint64_t> nmask((size_t)N, 0ull);
  std::vector<int32_t> ccol(C);
  int ncol = 0;
  for (int ci = 0; ci < C; ++ci) {
    uint64_t used = 0;
    for (int m = 0; m < kNsClElems; ++m) {
      const int e = cl_elem[(size_t)ci * kNsClElems + m];
      if (e >= 0)
        for (int a = 0; a < S; ++a) used |= nmask[node(e, a)];
    }
    if (used == ~0ull) return OFD_OK;  // past 64 colours: no clusters
    const int col = __builtin_ctzll(~used);
    ccol[ci] = col;
    ncol = std::max(ncol, col + 1);
    for (int m = 0; m < kNsClElems; ++m) {
      const int e = cl_elem[(size_t)ci * kNsClElems + m];
      if (e >= 0)
        for (int a = 0; a < S; ++a) nmask[node(e, a)] |= 1ull << col;
    }
  }
  std::vector<int32_t> coff(ncol + 1, 0), corder(C);
  for (int ci = 0; ci < C; ++ci) coff[ccol[ci] + 1]++;
  for (int k = 0; k < ncol; ++k) coff[k + 1] += coff[k];
  {
    std::vector<int32_t> pos(coff.begin(), coff.end() - 1);
    for (int ci = 0; ci < C; ++ci) corder[pos[ccol[ci]]++] = ci;
  }
  // the device tables in cluster-colour order; each slot's nonzero from the
  // pattern (row segments of h_col / h_row) and its first touch
  std::vector<int32_t> d_elem((size_t)C * kNsClElems), d_nslot(C);
  std::vector<uint16_t> d_map((size_t)C * kNsClElems * 64);
  std::vector<uint32_t> d_kf((size_t)C * kNsClSlots, 0x80000000u);
  std::vector<uint8_t> touched((size_t)c.icnt, 0);
  for (int r = 0; r < C; ++r) {
    const int ci = corder[r];
    for (int m = 0; m < kNsClElems; ++m) {
      const int e = cl_elem[(size_t)ci * kNsClElems + m];
      d_elem[(size_t)r * kNsClElems + m] = e >= 0 ? tpos[e] : -1;
    }
    std::copy(cl_map.begin() + (size_t)ci * kNsClElems * 64,
              cl_map.begin() + (size_t)(ci + 1) * kNsClElems * 64,
              d_map.begin() + (size_t)r * kNsClElems * 64);
    d_nslot[r] = cl_nslot[ci];
    for (int sl = 0; sl < cl_nslot[ci]; ++sl) {
      const auto pr = cl_pair[(size_t)ci * kNsClSlots + sl];
      int64_t k = c.h_col[pr.first];
      while (k < c.h_col[pr.first + 1] && c.h_row[k] != pr.second) ++k;
      OFD_CHECK(k < c.h_col[pr.first + 1], OFD_ERR_ARG,
                "cluster pair outside the sparsity pattern");
      d_kf[(size_t)r * kNsClSlots + sl] = (uint32_t)k | (touched[k] ? 0u : 0x80000000u);
      touched[k] = 1;
    }
  }
  int rc = dalloc(c, &c.d_cl_elem, d_elem.size());
  if (!rc) rc = dalloc(c, &c.d_cl_map, d_map.size());
  if (!rc) rc = dalloc(c, &c.d_cl_kf, d_kf.size());
  if (!rc) rc = dalloc(c, &c.d_cl_nslot, d_nslot.size());
  if (!rc) rc = h2d(c, c.d_cl_elem, d_elem.data(), d_elem.size() * sizeof(int32_t));
  if (!rc) rc = h2d(c, c.d_cl_map, d_map.data(), d_map.size() * sizeof(uint16_t));
  if (!rc) rc = h2d(c, c.d_cl_kf, d_kf.data(), d_kf.size() * sizeof(uint32_t));
  if (!rc) rc = h2d(c, c.d_cl_nslot, d_nslot.data(), d_nslot.size() * sizeof(int32_t));
  if (rc) return rc;
  c.cl_off = coff;
  return OFD_OK;
}

int ofd_build_pattern(ofd_ctx *x, int64_t *icnt_out) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  Ctx &c = x->c;
  int rc = need_mesh(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  const int N = c.nnode, E = c.nelem, S = c.nshl;
  // genSparStruc on the device (pattern.hip): incidence, rows, diagonal,
  // tmap and emap; the temporaries are freed on every path
  struct Tmp {
    std::vector<void *> p;
    ~Tmp() {
      for (void *q : p) (void)hipFree(q);
    }
    hipError_t get(void **out, size_t bytes) {
      hipError_t e = hipMalloc(out, bytes ? bytes : 1);
      if (e == hipSuccess) p.push_back(*out);
      return e;
    }
  } tmp;
  int32_t *d_cnt, *d_off, *d_pos, *d_inc, *d_err;
  OFD_HIP(tmp.get((void **)&d_cnt, sizeof(int32_t) * N));
  OFD_HIP(tmp.get((void **)&d_off, sizeof(int32_t) * ((size_t)N + 1)));
  OFD_HIP(tmp.get((void **)&d_pos, sizeof(int32_t) * N));
  OFD_HIP(tmp.get((void **)&d_inc, sizeof(int32_t) * (size_t)E * S));
  OFD_HIP(tmp.get((void **)&d_err, sizeof(int32_t)));
  OFD_HIP(hipMemsetAsync(d_cnt, 0, sizeof(int32_t) * N, c.stream));
  OFD_HIP(hipMemsetAsync(d_pos, 0, sizeof(int32_t) * N, c.stream));
  OFD_HIP(hipMemsetAsync(d_err, 0, sizeof(int32_t), c.stream));
  size_t scan_bytes = 0;
  OFD_HIP(scan_offsets(c, d_cnt, d_off, N, nullptr, scan_bytes));
  void *d_scan;
  OFD_HIP(tmp.get(&d_scan, scan_bytes));
  OFD_HIP(launch_node_elem_count(c, d_cnt));
  OFD_HIP(scan_offsets(c, d_cnt, d_off, N, d_scan, scan_bytes));
  OFD_HIP(launch_node_elem_fill(c, d_off, d_pos, d_inc));
  // row lengths, then col (genSparStruc.f90:27-45)
  OFD_HIP(launch_adjacency(c, d_off, d_inc, nullptr, d_cnt, nullptr, d_err));
  rc = dalloc(c, &c.d_col, (size_t)N + 1);
  if (rc) return rc;
  OFD_HIP(scan_offsets(c, d_cnt, c.d_col, N, d_scan, scan_bytes));
  int32_t err = 0;
  rc = d2h(c, &err, d_err, sizeof(int32_t));
  if (rc) return rc;
  OFD_CHECK(!(err & 1), OFD_ERR_ARG, "node adjacency exceeds 27*NSHL");
  c.h_col.assign(N + 1, 0);
  rc = d2h(c, c.h_col.data(), c.d_col, (N + 1) * sizeof(int32_t));
  if (rc) return rc;
  c.icnt = c.h_col[N];
  rc = dalloc(c, &c.d_row, (size_t)c.icnt);
  if (!rc) rc = dalloc(c, &c.d_diag, (size_t)N);
  if (!rc) rc = dalloc(c, &c.d_emap, (size_t)E * 64);
  if (!rc) rc = dalloc(c, &c.d_tmap, (size_t)c.icnt);
  if (rc) return rc;
  OFD_HIP(hipMemsetAsync(c.d_diag, 0xff, sizeof(int32_t) * N, c.stream));  // -1: none
  OFD_HIP(launch_adjacency(c, d_off, d_inc, c.d_col, c.d_row, c.d_diag, d_err));
  OFD_HIP(launch_tmap_emap(c, d_err));
  rc = d2h(c, &err, d_err, sizeof(int32_t));
  if (rc) return rc;
  OFD_CHECK(!(err & 2), OFD_ERR_ARG, "sparsity pattern is not symmetric");
  c.h_row.resize(c.icnt);
  rc = d2h(c, c.h_row.data(), c.d_row, c.icnt * sizeof(int32_t));
  if (rc) return rc;
  // greedy colouring: elements of one colour share no node.  Each node keeps
  // the colours of its elements as a bitset of nw 64-bit words, widened when
  // an element needs a colour past 64 nw (first fit needs at least as many
  // colours as the most elements at one node: 8 on a structured hex mesh,
  // tens on an unstructured tet mesh)
  int nw = 1;
  std::vector<uint64_t> nodemask((size_t)N, 0ull);
  std::vector<int32_t> ecol(E);
  std::vector<uint64_t> used;
  int ncolor = 0;
  for (int e = 0; e < E; ++e) {
    used.assign(nw, 0ull);
    for (int a = 0; a < S; ++a) {
      const uint64_t *m = &nodemask[(size_t)c.h_ien[(size_t)a * E + e] * nw];
      for (int w = 0; w < nw; ++w) used[w] |= m[w];
    }
    int col = 0;
    while (col < 64 * nw && ((used[col >> 6] >> (col & 63)) & 1ull)) ++col;
    if (col == 64 * nw) {
      std::vector<uint64_t> wider((size_t)N * (nw + 1), 0ull);
      for (int i = 0; i < N; ++i)
        for (int w = 0; w < nw; ++w) wider[(size_t)i * (nw + 1) + w] = nodemask[(size_t)i * nw + w];
      nodemask.swap(wider);
      ++nw;
    }
    ecol[e] = col;
    ncolor = std::max(ncolor, col + 1);
    for (int a = 0; a < S; ++a)
      nodemask[(size_t)c.h_ien[(size_t)a * E + e] * nw + (col >> 6)] |= 1ull << (col & 63);
  }
  c.color_off.assign(ncolor + 1, 0);
  for (int e = 0; e < E; ++e) c.color_off[ecol[e] + 1]++;
  c.h_elem_color = ecol;
  c.fe_dirty = true;
  for (int k = 0; k < ncolor; ++k) c.color_off[k + 1] += c.color_off[k];
  std::vector<int32_t> order(E);
  {
    std::vector<int32_t> pos(c.color_off.begin(), c.color_off.end() - 1);
    for (int e = 0; e < E; ++e) order[pos[ecol[e]]++] = e;
  }
  rc = dalloc(c, &c.d_color_elems, (size_t)E);
  if (!rc) rc = dalloc(c, &c.d_A64, (size_t)c.icnt * 16);
  if (rc) return rc;
  dfree(c.d_A32);
  dfree(c.d_Ap64);
  dfree(c.d_Ap32);
  dfree(c.t_A64);
  dfree(c.t_A32);
  dfree(c.t_Ap64);
  dfree(c.t_Ap32);
  c.t_mat_valid = c.t_prec_valid = c.t_a32_valid = c.t_ap32_valid = false;
  free_tem_transpose(c);
  rc = h2d(c, c.d_color_elems, order.data(), E * sizeof(int32_t));
  if (rc) return rc;
  {
    // node -> (position in colour order, node) entries in colour order, the
    // order in which a coloured scatter adds the elements' contributions
    const int nshl = c.nshl;
    std::vector<int32_t> ioff((size_t)N + 1, 0), inc((size_t)E * nshl);
    for (int e = 0; e < E; ++e)
      for (int a = 0; a < nshl; ++a) ioff[(size_t)c.h_ien[(size_t)a * E + e] + 1]++;
    for (int i = 0; i < N; ++i) ioff[i + 1] += ioff[i];
    std::vector<int32_t> pos(ioff.begin(), ioff.end() - 1);
    for (int t = 0; t < E; ++t) {
      const int e = order[t];
      for (int a = 0; a < nshl; ++a) inc[pos[c.h_ien[(size_t)a * E + e]]++] = t * 8 + a;
    }
    rc = dalloc(c, &c.d_inc_off, (size_t)N + 1);
    if (!rc) rc = dalloc(c, &c.d_inc, inc.size());
    if (!rc) rc = h2d(c, c.d_inc_off, ioff.data(), ioff.size() * sizeof(int32_t));
    if (!rc) rc = h2d(c, c.d_inc, inc.data(), inc.size() * sizeof(int32_t));
    if (rc) return rc;
  }
  // which element writes each nonzero first in the colour order: it stores,
  // the later ones add (the assemblies need no zeroing pass of the matrix)
  rc = dalloc(c, &c.d_first, (size_t)E);
  if (rc) return rc;
  {
    uint8_t *d_touched;
    OFD_HIP(tmp.get((void **)&d_touched, (size_t)c.icnt));
    OFD_HIP(hipMemsetAsync(d_touched, 0, (size_t)c.icnt, c.stream));
    OFD_HIP(launch_first_touch(c, d_touched));
  }
  if (c.nshl == 8 || c.nshl == 4 || c.nshl == 6) {
    rc = dalloc(c, &c.d_emap_c, (size_t)E * 64);
    if (!rc) rc = dalloc(c, &c.d_first_c, (size_t)E);
    if (!rc) rc = dalloc(c, &c.d_cpos, (size_t)E);
    if (rc) return rc;
    OFD_HIP(launch_colour_maps(c));
  }
  c.cl_off.clear();
  if (c.nshl == 8) {
    rc = build_ns_clusters(c, order);
    if (rc) return rc;
  }
  OFD_HIP(hipMemsetAsync(c.d_A64, 0, (size_t)c.icnt * 16 * sizeof(double), c.stream));
  OFD_HIP(hipStreamSynchronize(c.stream));
  c.mat_valid = c.prec_valid = c.a32_valid = c.ap32_valid = false;
  ++c.mat_version;
  if (icnt_out) *icnt_out = c.icnt;
  return OFD_OK;
}

int ofd_get_pattern(ofd_ctx *x, int32_t *col, int32_t *row) {
  OFD_CHECK(x && col && row, OFD_ERR_ARG, "null argument");
  Ctx &c = x->c;
  int rc = need_pattern(c);
  if (rc) return rc;
  for (int i = 0; i <= c.nnode; ++i) col[i] = c.h_col[i] + 1;
  for (int64_t k = 0; k < c.icnt; ++k) row[k] = c.h_row[k] + 1;
  return OFD_OK;
}

int ofd_check_pattern(ofd_ctx *x, int nnode, int64_t icnt, const int32_t *col,
                      const int32_t *row) {
  OFD_CHECK(x && col && row, OFD_ERR_ARG, "null argument");
  Ctx &c = x->c;
  int rc = need_pattern(c);
  if (rc) return rc;
  OFD_CHECK(nnode == c.nnode, OFD_ERR_ARG,
            "pattern check: NNODZu " + std::to_string(nnode) + " differs from the context's " +
                std::to_string(c.nnode));
  OFD_CHECK(icnt == c.icnt, OFD_ERR_ARG,
            "pattern check: icntu " + std::to_string(icnt) + " differs from the context's " +
                std::to_string(c.icnt));
  for (int i = 0; i <= nnode; ++i)
    OFD_CHECK(col[i] - 1 == c.h_col[i], OFD_ERR_ARG,
              "pattern check: col(" + std::to_string(i + 1) + ") differs from the context's");
  for (int64_t k = 0; k < icnt; ++k)
    OFD_CHECK(row[k] - 1 == c.h_row[k], OFD_ERR_ARG,
              "pattern check: row(" + std::to_string(k + 1) + ") differs from the context's");
  return OFD_OK;
}

int ofd_set_params(ofd_ctx *x, const double *prm) {
  OFD_CHECK(x && prm, OFD_ERR_ARG, "null argument");
  OFD_CHECK(prm[OFD_P_RHO] > 0.0 && prm[OFD_P_DELT] > 0.0, OFD_ERR_ARG,
            "rho and Delt must be positive");
  std::memcpy(x->c.prm, prm, sizeof(double) * OFD_NPARAMS);
  return OFD_OK;
}

int ofd_set_subdomains(ofd_ctx *x, int32_t id1, int32_t id2) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  x->c.elm_id_set[0] = id1;
  x->c.elm_id_set[1] = id2;
  return OFD_OK;
}

int ofd_set_ibc(ofd_ctx *x, const int32_t *ibc, int ncol) {
  OFD_CHECK(x && ibc, OFD_ERR_ARG, "null argument");
  OFD_CHECK(ncol >= 4 && ncol <= 8, OFD_ERR_ARG, "IBC needs 4 to 8 columns (u,v,w,p,..,T)");
  Ctx &c = x->c;
  int rc = need_mesh(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  std::vector<uint8_t> m(c.nnode, 0);
  for (int i = 0; i < c.nnode; ++i)
    for (int col = 0; col < ncol; ++col)
      if (ibc[(size_t)col * c.nnode + i] == 1) m[i] |= (uint8_t)(1u << col);
  return h2d(c, c.d_ibc, m.data(), m.size());
}

int ofd_set_bcvals(ofd_ctx *x, const double *vals) {
  OFD_CHECK(x && vals, OFD_ERR_ARG, "null argument");
  Ctx &c = x->c;
  int rc = need_mesh(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  return h2d(c, c.d_bcval, vals, (size_t)c.nnode * 3 * sizeof(double));
}

int ofd_apply_dirichlet(ofd_ctx *x) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  Ctx &c = x->c;
  int rc = need_mesh(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  OFD_HIP(launch_apply_dirichlet(c));
  OFD_HIP(hipStreamSynchronize(c.stream));
  return OFD_OK;
}

// SOLU / SOLP (/ SOLPHI) are views of the last solved system's increment
static hipError_t sol_view(Ctx &c) {
  if (c.sol_src == 2 && c.d_xV) return launch_nsvof_to_soa(c, c.d_xV);
  return launch_node4_to_soa(c, c.d_x, c.d_field[OFD_SOLU], c.d_field[OFD_SOLP]);
}

int ofd_set_field(ofd_ctx *x, int field, const double *host) {
  OFD_CHECK(x && host, OFD_ERR_ARG, "null argument");
  OFD_CHECK(field >= 0 && field < OFD_NFIELDS, OFD_ERR_ARG, "unknown field");
  Ctx &c = x->c;
  int rc = need_mesh(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  rc = h2d(c, c.d_field[field], host, (size_t)field_len(c, field) * sizeof(double));
  if (rc) return rc;
  if (field == OFD_MUSCALE) c.mus_on = true;
  if (field == OFD_KSCALE) c.ks_on = true;
  if (field == OFD_SOLU || field == OFD_SOLP) {
    c.sol_src = 0;
    OFD_HIP(launch_soa_to_node4(c, c.d_field[OFD_SOLU], c.d_field[OFD_SOLP], c.d_x));
    OFD_HIP(hipStreamSynchronize(c.stream));
  }
  return OFD_OK;
}

int ofd_get_field(ofd_ctx *x, int field, double *host) {
  OFD_CHECK(x && host, OFD_ERR_ARG, "null argument");
  OFD_CHECK(field >= 0 && field < OFD_NFIELDS, OFD_ERR_ARG, "unknown field");
  Ctx &c = x->c;
  int rc = need_mesh(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  if (field == OFD_SOLU || field == OFD_SOLP || field == OFD_SOLPHI) OFD_HIP(sol_view(c));
  return d2h(c, host, c.d_field[field], (size_t)field_len(c, field) * sizeof(double));
}

// the elements holding registered weak faces, grouped by element colour,
// with their faces as a bitmask (bit fo-1), on the device
static int ensure_faces(Ctx &c) {
  OFD_CHECK(c.h_face_e.empty() || c.nshl == 8 || c.nshl == 4, OFD_ERR_UNSUPPORTED,
            "boundary faces are implemented for trilinear hexes and linear tets");
  if (!c.fe_dirty) return OFD_OK;
  c.fe_color_off.clear();
  if (c.h_face_e.empty()) {
    c.fe_dirty = false;
    return OFD_OK;
  }
  const int E = c.nelem, ncolor = (int)c.color_off.size() - 1;
  OFD_CHECK((int)c.h_elem_color.size() == E, OFD_ERR_ARG, "no element colouring");
  std::vector<int32_t> mask(E, 0);
  for (size_t f = 0; f < c.h_face_e.size(); ++f)
    mask[c.h_face_e[f]] |= 1 << (c.h_face_or[f] - 1);
  c.fe_color_off.assign(ncolor + 1, 0);
  for (int e = 0; e < E; ++e)
    if (mask[e]) c.fe_color_off[c.h_elem_color[e] + 1]++;
  for (int k = 0; k < ncolor; ++k) c.fe_color_off[k + 1] += c.fe_color_off[k];
  const int nfe = c.fe_color_off[ncolor];
  std::vector<int32_t> el(nfe), mk(nfe);
  std::vector<int32_t> pos(c.fe_color_off.begin(), c.fe_color_off.end() - 1);
  for (int e = 0; e < E; ++e)
    if (mask[e]) {
      const int p = pos[c.h_elem_color[e]]++;
      el[p] = e;
      mk[p] = mask[e];
    }
  int rc = dalloc(c, &c.d_fe_elem, (size_t)nfe);
  if (!rc) rc = dalloc(c, &c.d_fe_mask, (size_t)nfe);
  if (!rc) rc = h2d(c, c.d_fe_elem, el.data(), nfe * sizeof(int32_t));
  if (!rc) rc = h2d(c, c.d_fe_mask, mk.data(), nfe * sizeof(int32_t));
  if (rc) return rc;
  c.fe_dirty = false;
  return OFD_OK;
}

int ofd_set_weak_faces(ofd_ctx *x, int nface, const int32_t *f2e, const int32_t *face_or) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  OFD_CHECK(nface >= 0, OFD_ERR_ARG, "nface < 0");
  Ctx &c = x->c;
  int rc = need_mesh(c);
  if (rc) return rc;
  OFD_CHECK(nface == 0 || (f2e && face_or), OFD_ERR_ARG, "null argument");
  OFD_CHECK(nface == 0 || c.nshl == 8 || c.nshl == 4, OFD_ERR_UNSUPPORTED,
            "boundary faces are implemented for trilinear hexes and linear tets");
  const int nfo = c.nshl == 8 ? 6 : 4;
  std::vector<int32_t> fe(nface), fo(nface);
  std::vector<uint8_t> seen((size_t)c.nelem, 0);
  for (int f = 0; f < nface; ++f) {
    const int e = f2e[f] - 1, o = face_or[f];
    OFD_CHECK(e >= 0 && e < c.nelem, OFD_ERR_ARG, "F2E out of range [1, NELEM]");
    OFD_CHECK(o >= 1 && o <= nfo, OFD_ERR_ARG,
              nfo == 6 ? "FACE_OR out of range [1, 6]" : "FACE_OR out of range [1, 4]");
    OFD_CHECK(!((seen[e] >> (o - 1)) & 1), OFD_ERR_ARG, "face listed twice");
    seen[e] |= (uint8_t)(1u << (o - 1));
    fe[f] = e;
    fo[f] = o;
  }
  c.h_face_e.swap(fe);
  c.h_face_or.swap(fo);
  c.fe_dirty = true;
  return OFD_OK;
}

int ofd_assemble_ns(ofd_ctx *x, int tensor_flag, int subdomain_flag, double *cfl) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  Ctx &c = x->c;
  int rc = need_pattern(c);
  if (rc) return rc;
  OFD_CHECK(c.nshl == 8 || c.nshl == 4 || c.nshl == 6, OFD_ERR_UNSUPPORTED,
            "the GPU NS assembly is implemented for trilinear hexes, linear tets and prisms");
  OFD_CHECK((tensor_flag & ~(OFD_ASSEMBLE_VEC | OFD_ASSEMBLE_MAT)) == 0, OFD_ERR_ARG,
            "tensor_flag must combine OFD_ASSEMBLE_VEC / OFD_ASSEMBLE_MAT");
  OFD_HIP(hipSetDevice(c.device));
  rc = ensure_pts(c);
  if (rc) return rc;
  OFD_HIP(launch_alpha_stage(c));
  if (tensor_flag & OFD_ASSEMBLE_MAT) {
    // no zeroing: the first element to reach a nonzero stores into it
    c.mat_valid = true;
    c.prec_valid = c.a32_valid = c.ap32_valid = false;
    ++c.mat_version;
  }
  OFD_HIP(hipMemsetAsync(c.d_negdet, 0, sizeof(int32_t), c.stream));
  OFD_HIP(launch_assemble_ns(c, tensor_flag, subdomain_flag));
  // FaceAssembly_NS_weak and FaceAssembly_NS_outflow (assembleNavSto.f90:54-60)
  rc = ensure_faces(c);
  if (rc) return rc;
  for (int col = 0; col + 1 < (int)c.fe_color_off.size(); ++col)
    OFD_HIP(launch_face_ns(c, tensor_flag, subdomain_flag, col));
  double *d_out = c.d_part;  // scratch for the CFL maximum
  OFD_HIP(launch_max_reduce(c, c.d_cfl, c.nelem, d_out));
  double h[1] = {0.0};
  int32_t neg = 0;
  OFD_HIP(hipMemcpyAsync(h, d_out, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  OFD_HIP(hipMemcpyAsync(&neg, c.d_negdet, sizeof(int32_t), hipMemcpyDeviceToHost,
                         c.stream));
  OFD_HIP(hipStreamSynchronize(c.stream));
  if (cfl) *cfl = h[0];
  if (neg > 0) {
    set_error("negative determinant Jacobian at " + std::to_string(neg) +
              " Gauss points (eval_shape_fem warning)");
    return OFD_ERR_NEGDETJ;
  }
  return OFD_OK;
}

int ofd_get_rhs(ofd_ctx *x, double *rhsu, double *rhsp) {
  OFD_CHECK(x && rhsu && rhsp, OFD_ERR_ARG, "null argument");
  Ctx &c = x->c;
  int rc = need_mesh(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  OFD_HIP(launch_node4_to_soa(c, c.d_rhs, c.d_field[OFD_SOLU], c.d_field[OFD_SOLP]));
  rc = d2h(c, rhsu, c.d_field[OFD_SOLU], (size_t)c.nnode * 3 * sizeof(double));
  if (!rc) rc = d2h(c, rhsp, c.d_field[OFD_SOLP], (size_t)c.nnode * sizeof(double));
  // restore the solution view of the scratch fields
  OFD_HIP(sol_view(c));
  OFD_HIP(hipStreamSynchronize(c.stream));
  return rc;
}

int ofd_set_rhs(ofd_ctx *x, const double *rhsu, const double *rhsp) {
  OFD_CHECK(x && rhsu && rhsp, OFD_ERR_ARG, "null argument");
  Ctx &c = x->c;
  int rc = need_mesh(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  rc = h2d(c, c.d_r, rhsu, (size_t)c.nnode * 3 * sizeof(double));
  if (rc) return rc;
  std::vector<double> tmp;  // p goes through the b buffer
  rc = h2d(c, c.d_b, rhsp, (size_t)c.nnode * sizeof(double));
  if (rc) return rc;
  OFD_HIP(launch_soa_to_node4(c, c.d_r, c.d_b, c.d_rhs));
  OFD_HIP(hipStreamSynchronize(c.stream));
  return OFD_OK;
}

int ofd_get_lhs(ofd_ctx *x, double *k11, double *g, double *d1, double *m) {
  OFD_CHECK(x && k11 && g && d1 && m, OFD_ERR_ARG, "null argument");
  Ctx &c = x->c;
  int rc = need_pattern(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  std::vector<double> A((size_t)c.icnt * 16);
  rc = d2h(c, A.data(), c.d_A64, A.size() * sizeof(double));
  if (rc) return rc;
  for (int64_t k = 0; k < c.icnt; ++k) {
    const double *b = &A[(size_t)k * 16];
    for (int r = 0; r < 3; ++r) {
      for (int cc = 0; cc < 3; ++cc) k11[(size_t)k * 9 + r * 3 + cc] = b[r * 4 + cc];
      g[(size_t)k * 3 + r] = b[r * 4 + 3];
      d1[(size_t)k * 3 + r] = b[12 + r];
    }
    m[k] = b[15];
  }
  return OFD_OK;
}

int ofd_set_lhs(ofd_ctx *x, const double *k11, const double *g, const double *d1,
                const double *m) {
  OFD_CHECK(x && k11 && g && d1 && m, OFD_ERR_ARG, "null argument");
  Ctx &c = x->c;
  int rc = need_pattern(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  std::vector<double> A((size_t)c.icnt * 16);
  for (int64_t k = 0; k < c.icnt; ++k) {
    double *b = &A[(size_t)k * 16];
    for (int r = 0; r < 3; ++r) {
      for (int cc = 0; cc < 3; ++cc) b[r * 4 + cc] = k11[(size_t)k * 9 + r * 3 + cc];
      b[r * 4 + 3] = g[(size_t)k * 3 + r];
      b[12 + r] = d1[(size_t)k * 3 + r];
    }
    b[15] = m[k];
  }
  rc = h2d(c, c.d_A64, A.data(), A.size() * sizeof(double));
  if (rc) return rc;
  c.mat_valid = true;
  c.prec_valid = c.a32_valid = c.ap32_valid = false;
  ++c.mat_version;
  return OFD_OK;
}

int ofd_spmv_up(ofd_ctx *x, int precision, const double *xu, const double *xp,
                double *yu, double *yp) {
  OFD_CHECK(x && xu && xp && yu && yp, OFD_ERR_ARG, "null argument");
  OFD_CHECK(precision == OFD_FP64 || precision == OFD_FP32, OFD_ERR_ARG, "bad precision");
  Ctx &c = x->c;
  int rc = need_pattern(c);
  if (rc) return rc;
  OFD_CHECK(c.mat_valid, OFD_ERR_ARG, "no assembled matrix");
  OFD_HIP(hipSetDevice(c.device));
  if (precision == OFD_FP32) {
    rc = ensure_a32(c);
    if (rc) return rc;
  }
  double *su = c.d_field[OFD_SOLU], *sp = c.d_field[OFD_SOLP];
  rc = h2d(c, su, xu, (size_t)c.nnode * 3 * sizeof(double));
  if (!rc) rc = h2d(c, sp, xp, (size_t)c.nnode * sizeof(double));
  if (rc) return rc;
  OFD_HIP(launch_soa_to_node4(c, su, sp, c.d_b));
  OFD_HIP(launch_spmv(c, precision, c.d_b, c.d_r, nullptr, 0));
  OFD_HIP(launch_node4_to_soa(c, c.d_r, su, sp));
  rc = d2h(c, yu, su, (size_t)c.nnode * 3 * sizeof(double));
  if (!rc) rc = d2h(c, yp, sp, (size_t)c.nnode * sizeof(double));
  OFD_HIP(sol_view(c));
  OFD_HIP(hipStreamSynchronize(c.stream));
  return rc;
}

int ofd_gmres_up(ofd_ctx *x, int precision, int ortho, double tol, int kspace,
                 int kspace_min, double *solu, double *solp, int *iters, double *resid) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  OFD_CHECK(precision == OFD_FP64 || precision == OFD_FP32, OFD_ERR_ARG, "bad precision");
  OFD_CHECK(ortho == OFD_ORTHO_MGS || ortho == OFD_ORTHO_CGS2 || ortho == OFD_ORTHO_CGS1,
            OFD_ERR_ARG, "bad ortho");
  Ctx &c = x->c;
  int rc = need_pattern(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  OFD_HIP(hipMemsetAsync(c.d_x, 0, (size_t)c.nnode * 4 * sizeof(double), c.stream));
  double bn = 0.0;
  rc = norm2_dev(c, c.d_rhs, &bn);
  if (rc) return rc;
  OFD_CHECK(std::isfinite(bn), OFD_ERR_NONFINITE, "SparseGMRES_up: non-finite right-hand side");
  OFD_HIP(launch_outer_reset(c, 0.0));
  int it = 0;
  double res = 0.0;
  rc = gmres_cycle(c, precision, ortho, c.d_rhs, tol, kspace, kspace_min, c.d_x, &it,
                   &res, nullptr);
  if (iters) *iters = it;
  if (resid) *resid = res;
  if (rc) return rc;
  c.sol_src = 0;
  OFD_HIP(launch_node4_to_soa(c, c.d_x, c.d_field[OFD_SOLU], c.d_field[OFD_SOLP]));
  if (solu) rc = d2h(c, solu, c.d_field[OFD_SOLU], (size_t)c.nnode * 3 * sizeof(double));
  if (!rc && solp) rc = d2h(c, solp, c.d_field[OFD_SOLP], (size_t)c.nnode * sizeof(double));
  OFD_HIP(hipStreamSynchronize(c.stream));
  return rc;
}

// Restarted GMRES with an FP64 outer residual on the current system (NS or,
// under TemSystem, T): x (gm_n 4-tuples) starts at 0 and solves A x = b.
static int solve_core(Ctx &c, int precision, int ortho, const double *b, double *x,
                      double *r, double rtol, double inner_tol, int kspace, int max_cycles,
                      int *iters, int *cycles, double *relres) {
  OFD_HIP(hipMemsetAsync(x, 0, (size_t)c.gm_n * 4 * sizeof(double), c.stream));
  double bnorm = 0.0;
  int rc = norm2_dev(c, b, &bnorm);
  if (rc) return rc;
  int total = 0, cyc = 0;
  double rel = 0.0;
  if (!(bnorm > 0.0)) {
    if (iters) *iters = 0;
    if (cycles) *cycles = 0;
    if (relres) *relres = 0.0;
    return std::isfinite(bnorm) ? OFD_OK : OFD_ERR_NONFINITE;
  }
  const double *rhs = b;
  c.solve_hist.clear();
  // Cycle c+1 is queued before the host reads cycle c's FP64 residual, so
  // the device never waits for the host between cycles.  The residual norm
  // kernel sets GmresState::outer_done once |b - A x| <= rtol |b|; every
  // kernel of a cycle queued behind that one returns at once (x and r stay
  // as they are).  Each cycle's state is copied to its own pinned slot.
  OFD_HIP(launch_outer_reset(c, rtol * bnorm));
  int done_cycles = 0;
  int32_t seq[2] = {0, 0};
  for (cyc = 1; cyc <= max_cycles + 1; ++cyc) {
    if (cyc <= max_cycles) {
      rc = gmres_cycle(c, precision, ortho, rhs, inner_tol, kspace, 0, x, nullptr, nullptr,
                       nullptr, false, cyc & 1);
      if (rc) break;
      // FP64 outer residual r = b - A x and its norm
      if (c.gm_kind == 1)
        OFD_HIP(launch_spmv_s(c, OFD_FP64, x, r, b, 1));
      else if (c.gm_kind == 2)
        OFD_HIP(launch_spmv_v(c, OFD_FP64, x, r, b, 1));
      else
        OFD_HIP(launch_spmv(c, OFD_FP64, x, r, b, 1));
      seq[cyc & 1] = ++c.rec_seq;
      OFD_HIP(launch_norm2(c, r, 0, 1, &c.h_rec[cyc & 1], seq[cyc & 1]));
      rhs = r;
    }
    if (cyc >= 2) {
      const int p = cyc - 1;  // the cycle whose result is read now
      CycleRecord st;
      rc = wait_record(c, p & 1, seq[p & 1], &st);
      if (rc) break;
      done_cycles = p;
      total += st.iters;
      if (!std::isfinite(st.resid) || !std::isfinite(st.rnorm0)) {
        rc = OFD_ERR_NONFINITE;
        set_error("GMRES: non-finite residual");
        break;
      }
      rel = st.rnorm / bnorm;
      c.solve_hist.push_back(rel);
      if (!std::isfinite(rel)) {
        rc = OFD_ERR_NONFINITE;
        set_error("solve: non-finite residual");
        break;
      }
      if (rel <= rtol) break;
    }
  }
  // the queued cycle behind the last one read is a no-op or its result is
  // discarded with the error; drain it before touching x
  OFD_HIP(hipStreamSynchronize(c.stream));
  const int prc = probe_harvest(c);
  if (!rc) rc = prc;
  if (iters) *iters = total;
  if (cycles) *cycles = done_cycles;
  if (relres) *relres = rel;
  if (rc) return rc;
  if (rel > rtol) {
    set_error("solve: tolerance not reached within max_cycles");
    return OFD_ERR_NOCONV;
  }
  return OFD_OK;
}

int ofd_solve(ofd_ctx *x, int precision, int ortho, double rtol, double inner_tol,
              int kspace, int max_cycles, int *iters, int *cycles, double *relres) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  OFD_CHECK(precision == OFD_FP64 || precision == OFD_FP32, OFD_ERR_ARG, "bad precision");
  OFD_CHECK(max_cycles >= 1, OFD_ERR_ARG, "max_cycles >= 1");
  Ctx &c = x->c;
  int rc = need_pattern(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  rc = solve_core(c, precision, ortho, c.d_rhs, c.d_x, c.d_r, rtol, inner_tol, kspace,
                  max_cycles, iters, cycles, relres);
  if (rc && rc != OFD_ERR_NOCONV) return rc;
  c.sol_src = 0;
  OFD_HIP(launch_node4_to_soa(c, c.d_x, c.d_field[OFD_SOLU], c.d_field[OFD_SOLP]));
  OFD_HIP(hipStreamSynchronize(c.stream));
  return rc;
}

int ofd_solve_history(ofd_ctx *x, double *relres, int maxn, int *n) {
  OFD_CHECK(x && n, OFD_ERR_ARG, "null argument");
  const std::vector<double> &h = x->c.solve_hist;
  *n = (int)h.size();
  for (int i = 0; i < maxn && i < (int)h.size(); ++i) relres[i] = h[i];
  return OFD_OK;
}

int ofd_newton_update(ofd_ctx *x) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  Ctx &c = x->c;
  int rc = need_mesh(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  OFD_HIP(launch_axpy_fields(c, 0));
  OFD_HIP(hipStreamSynchronize(c.stream));
  return OFD_OK;
}

int ofd_predictor(ofd_ctx *x) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  Ctx &c = x->c;
  int rc = need_mesh(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  OFD_HIP(launch_axpy_fields(c, 1));
  if (c.quench) {
    OFD_HIP(launch_fields_s(c, 1, nullptr, OFD_TG, OFD_TGOLD, OFD_RTG, OFD_RTGOLD));
    OFD_HIP(launch_fields_s(c, 1, nullptr, OFD_PHIG, OFD_PHIGOLD, OFD_RPHIG,
                            OFD_RPHIGOLD));
  }
  OFD_HIP(hipStreamSynchronize(c.stream));
  return OFD_OK;
}

int ofd_advance(ofd_ctx *x) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  Ctx &c = x->c;
  int rc = need_mesh(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  OFD_HIP(launch_axpy_fields(c, 2));
  if (c.quench) {
    OFD_HIP(launch_fields_s(c, 2, nullptr, OFD_TG, OFD_TGOLD, OFD_RTG, OFD_RTGOLD));
    OFD_HIP(launch_fields_s(c, 2, nullptr, OFD_PHIG, OFD_PHIGOLD, OFD_RPHIG,
                            OFD_RPHIGOLD));
  }
  OFD_HIP(hipStreamSynchronize(c.stream));
  return OFD_OK;
}

int ofd_rhs_norms(ofd_ctx *x, double *norms2) {
  OFD_CHECK(x && norms2, OFD_ERR_ARG, "null argument");
  Ctx &c = x->c;
  int rc = need_mesh(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  // two-pass fixed-order reduction on the device into GmresState::scal[1..2]
  OFD_HIP(launch_norm2_up(c, c.d_rhs, 1));
  OFD_HIP(hipMemcpyAsync(c.h_gs, c.d_gs, sizeof(GmresState), hipMemcpyDeviceToHost,
                         c.stream));
  OFD_HIP(hipStreamSynchronize(c.stream));
  norms2[0] = c.h_gs->scal[1];
  norms2[1] = c.h_gs->scal[2];
  return OFD_OK;
}

int ofd_time_spmv(ofd_ctx *x, int precision, int reps, double *ms) {
  OFD_CHECK(x && ms && reps > 0, OFD_ERR_ARG, "bad argument");
  Ctx &c = x->c;
  int rc = need_pattern(c);
  if (rc) return rc;
  OFD_CHECK(c.mat_valid, OFD_ERR_ARG, "no assembled matrix");
  OFD_HIP(hipSetDevice(c.device));
  if (precision == OFD_FP32) {
    rc = ensure_a32(c);
    if (rc) return rc;
  }
  hipEvent_t e0, e1;
  OFD_HIP(hipEventCreate(&e0));
  OFD_HIP(hipEventCreate(&e1));
  OFD_HIP(launch_spmv(c, precision, c.d_x, c.d_r, nullptr, 0));  // warm
  OFD_HIP(hipEventRecord(e0, c.stream));
  for (int r = 0; r < reps; ++r) OFD_HIP(launch_spmv(c, precision, c.d_x, c.d_r, nullptr, 0));
  OFD_HIP(hipEventRecord(e1, c.stream));
  OFD_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  OFD_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = (double)t / reps;
  return OFD_OK;
}

int ofd_time_faces(ofd_ctx *x, int tensor_flag, int reps, double *ms) {
  OFD_CHECK(x && ms && reps > 0, OFD_ERR_ARG, "bad argument");
  Ctx &c = x->c;
  int rc = need_pattern(c);
  if (rc) return rc;
  rc = ensure_faces(c);
  if (rc) return rc;
  OFD_CHECK(!c.fe_color_off.empty(), OFD_ERR_ARG, "no weak faces (ofd_set_weak_faces)");
  OFD_HIP(hipSetDevice(c.device));
  OFD_HIP(launch_alpha_stage(c));
  hipEvent_t e0, e1;
  OFD_HIP(hipEventCreate(&e0));
  OFD_HIP(hipEventCreate(&e1));
  OFD_HIP(hipEventRecord(e0, c.stream));
  for (int r = 0; r < reps; ++r)
    for (int col = 0; col + 1 < (int)c.fe_color_off.size(); ++col)
      OFD_HIP(launch_face_ns(c, tensor_flag, 3, col));
  OFD_HIP(hipEventRecord(e1, c.stream));
  OFD_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  OFD_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // the launches added into the system: it no longer holds an assembly
  c.mat_valid = false;
  c.prec_valid = c.a32_valid = c.ap32_valid = false;
  ++c.mat_version;
  *ms = (double)t / reps;
  return OFD_OK;
}

int ofd_time_assemble(ofd_ctx *x, int tensor_flag, int reps, double *ms) {
  OFD_CHECK(x && ms && reps > 0, OFD_ERR_ARG, "bad argument");
  Ctx &c = x->c;
  int rc = need_pattern(c);
  if (rc) return rc;
  rc = ensure_faces(c);
  if (rc) return rc;
  OFD_CHECK(c.nshl == 8 || c.nshl == 4 || c.nshl == 6, OFD_ERR_UNSUPPORTED,
            "the GPU NS assembly is implemented for trilinear hexes, linear tets and prisms");
  OFD_HIP(hipSetDevice(c.device));
  rc = ensure_pts(c);
  if (rc) return rc;
  hipEvent_t e0, e1;
  OFD_HIP(hipEventCreate(&e0));
  OFD_HIP(hipEventCreate(&e1));
  OFD_HIP(hipEventRecord(e0, c.stream));
  for (int r = 0; r < reps; ++r) {
    OFD_HIP(launch_alpha_stage(c));
    OFD_HIP(hipMemsetAsync(c.d_negdet, 0, sizeof(int32_t), c.stream));
    OFD_HIP(launch_assemble_ns(c, tensor_flag, 3));
    for (int col = 0; col + 1 < (int)c.fe_color_off.size(); ++col)
      OFD_HIP(launch_face_ns(c, tensor_flag, 3, col));
    OFD_HIP(launch_max_reduce(c, c.d_cfl, c.nelem, c.d_part));
  }
  OFD_HIP(hipEventRecord(e1, c.stream));
  OFD_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  OFD_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (tensor_flag & OFD_ASSEMBLE_MAT) {
    c.mat_valid = true;
    c.prec_valid = c.a32_valid = c.ap32_valid = false;
    ++c.mat_version;
  }
  *ms = (double)t / reps;
  return OFD_OK;
}

static int ensure_tem(Ctx &c);
static int ensure_aT32(Ctx &c);
static int ensure_nsvof(Ctx &c);

int ofd_time_quench(ofd_ctx *x, int what, int arg, int reps, double *ms) {
  OFD_CHECK(x && ms && reps > 0, OFD_ERR_ARG, "bad argument");
  OFD_CHECK(what >= OFD_TIME_ASM_NSVOF && what <= OFD_TIME_SPMV_TEM, OFD_ERR_ARG,
            "unknown kernel");
  Ctx &c = x->c;
  const bool nsvof = what == OFD_TIME_ASM_NSVOF || what == OFD_TIME_SPMV_NSVOF;
  const bool spmv = what >= OFD_TIME_SPMV_NSVOF;
  int rc = nsvof ? ensure_nsvof(c) : ensure_tem(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  if (!spmv) {
    OFD_CHECK(c.quench, OFD_ERR_ARG, "no quench parameters (ofd_set_quench_params)");
    OFD_CHECK(c.nshl == 8 || c.nshl == 4 || c.nshl == 6, OFD_ERR_UNSUPPORTED,
              "the GPU element assembly is implemented for trilinear hexes, linear tets and "
              "prisms");
    OFD_CHECK((arg & ~(OFD_ASSEMBLE_VEC | OFD_ASSEMBLE_MAT)) == 0, OFD_ERR_ARG,
              "arg must combine OFD_ASSEMBLE_VEC / OFD_ASSEMBLE_MAT");
    rc = ensure_pts(c);
    if (rc) return rc;
  } else {
    OFD_CHECK(arg == OFD_FP64 || arg == OFD_FP32, OFD_ERR_ARG, "bad precision");
    OFD_CHECK(nsvof ? c.matV_valid : c.matT_valid, OFD_ERR_ARG, "no assembled system");
    if (arg == OFD_FP32) {
      rc = nsvof ? ensure_aV32(c) : ensure_aT32(c);
      if (rc) return rc;
    }
  }
  hipEvent_t e0, e1;
  OFD_HIP(hipEventCreate(&e0));
  OFD_HIP(hipEventCreate(&e1));
  const int ncol = (int)c.color_off.size() - 1;
  for (int r = -1; r < reps; ++r) {  // r = -1: warm-up launch
    if (r == 0) OFD_HIP(hipEventRecord(e0, c.stream));
    if (what == OFD_TIME_ASM_NSVOF) {
      OFD_HIP(launch_alpha_stage(c, true));
      if (arg & OFD_ASSEMBLE_VEC) OFD_HIP(launch_zero(c, c.d_rhsV, (size_t)c.nsvV_n4 * 4));
      for (int col = 0; col < ncol; ++col) OFD_HIP(launch_assemble_nsvof_hex(c, arg, col));
    } else if (what == OFD_TIME_ASM_TEM) {
      OFD_HIP(launch_alpha_stage(c, false));
      if (arg & OFD_ASSEMBLE_VEC) OFD_HIP(launch_zero(c, c.d_rhsT, (size_t)c.temT_n4 * 4));
      OFD_HIP(launch_assemble_tem(c, arg));
    } else if (what == OFD_TIME_SPMV_NSVOF) {
      NsvofSystem sys(c);
      OFD_HIP(launch_spmv_v(c, arg, c.d_xV, c.d_yV, nullptr, 0));
    } else {
      TemSystem sys(c);
      OFD_HIP(launch_spmv_s(c, arg, c.d_xT, c.d_yT, nullptr, 0));
    }
  }
  OFD_HIP(hipEventRecord(e1, c.stream));
  OFD_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  OFD_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (!spmv && (arg & OFD_ASSEMBLE_MAT)) {
    if (nsvof) {
      c.matV_valid = true;
      c.precV_valid = c.pV_valid = c.aV32_valid = c.apV32_valid = false;
      c.t_matV_valid = false;
    } else {
      c.matT_valid = true;
      c.precT_valid = c.aT32_valid = c.apT32_valid = false;
      ++c.matT_version;
    }
  }
  *ms = (double)t / reps;
  return OFD_OK;
}

// ---- temperature system ----

// LHSTem for the current pattern and the zero-padded T vectors
static int ensure_tem(Ctx &c) {
  int rc = need_pattern(c);
  if (rc) return rc;
  const int n4 = (c.nnode + 3) / 4;
  if (c.temT_icnt != c.icnt) {
    rc = dalloc(c, &c.d_AT, (size_t)c.icnt);
    if (!rc) rc = dalloc(c, &c.d_AT32, (size_t)0);
    if (!rc) rc = dalloc(c, &c.d_ApT, (size_t)0);
    if (!rc) rc = dalloc(c, &c.d_ApT32, (size_t)0);
    if (rc) return rc;
    OFD_HIP(hipMemsetAsync(c.d_AT, 0, (size_t)c.icnt * sizeof(double), c.stream));
    c.temT_icnt = c.icnt;
    c.matT_valid = c.precT_valid = c.aT32_valid = c.apT32_valid = false;
    free_tem_transpose(c);
    ++c.matT_version;
  }
  if (c.temT_n4 != n4) {
    const size_t np = (size_t)n4 * 4;
    rc = dalloc(c, &c.d_rhsT, np);
    if (!rc) rc = dalloc(c, &c.d_xT, np);
    if (!rc) rc = dalloc(c, &c.d_rT, np);
    if (!rc) rc = dalloc(c, &c.d_yT, np);
    if (!rc) rc = dalloc(c, &c.d_PT, (size_t)c.nnode);
    if (rc) return rc;
    OFD_HIP(hipMemsetAsync(c.d_rhsT, 0, np * sizeof(double), c.stream));
    OFD_HIP(hipMemsetAsync(c.d_xT, 0, np * sizeof(double), c.stream));
    OFD_HIP(hipMemsetAsync(c.d_rT, 0, np * sizeof(double), c.stream));
    OFD_HIP(hipMemsetAsync(c.d_yT, 0, np * sizeof(double), c.stream));
    c.temT_n4 = n4;
  }
  return OFD_OK;
}

static int ensure_aT32(Ctx &c) {
  OFD_CHECK(c.matT_valid, OFD_ERR_ARG, "no assembled temperature matrix");
  if (c.aT32_valid) return OFD_OK;
  if (!c.d_AT32) {
    int rc = dalloc(c, &c.d_AT32, (size_t)c.icnt);
    if (rc) return rc;
  }
  OFD_HIP(launch_convert_aT32(c));
  c.aT32_valid = true;
  return OFD_OK;
}

int ofd_set_quench_params(ofd_ctx *x, const double *q) {
  OFD_CHECK(x && q, OFD_ERR_ARG, "null argument");
  Ctx &c = x->c;
  for (int i = 0; i < OFD_NQPARAMS; ++i) c.qprm[i] = q[i];
  OFD_CHECK(c.qprm[OFD_Q_TS] != 0.0, OFD_ERR_ARG, "T_SAT must be nonzero");
  c.quench = true;
  return OFD_OK;
}

int ofd_assemble_tem(ofd_ctx *x, int tensor_flag) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  Ctx &c = x->c;
  int rc = ensure_tem(c);
  if (rc) return rc;
  OFD_CHECK(c.quench, OFD_ERR_ARG, "no quench parameters (ofd_set_quench_params)");
  OFD_CHECK(c.nshl == 8 || c.nshl == 4 || c.nshl == 6, OFD_ERR_UNSUPPORTED,
            "the GPU element assembly is implemented for trilinear hexes, linear tets and "
            "prisms");
  OFD_CHECK((tensor_flag & ~(OFD_ASSEMBLE_VEC | OFD_ASSEMBLE_MAT)) == 0, OFD_ERR_ARG,
            "tensor_flag must combine OFD_ASSEMBLE_VEC / OFD_ASSEMBLE_MAT");
  OFD_HIP(hipSetDevice(c.device));
  rc = ensure_pts(c);
  if (rc) return rc;
  OFD_HIP(launch_alpha_stage(c));
  if (tensor_flag & OFD_ASSEMBLE_VEC)
    OFD_HIP(launch_zero(c, c.d_rhsT, (size_t)c.temT_n4 * 4));
  if (tensor_flag & OFD_ASSEMBLE_MAT) {
    // no zeroing: the first element to reach a nonzero stores into it
    c.matT_valid = true;
    c.precT_valid = c.aT32_valid = c.apT32_valid = false;
    ++c.matT_version;
  }
  OFD_HIP(hipMemsetAsync(c.d_negdet, 0, sizeof(int32_t), c.stream));
  OFD_HIP(launch_assemble_tem(c, tensor_flag));
  int32_t neg = 0;
  OFD_HIP(hipMemcpyAsync(&neg, c.d_negdet, sizeof(int32_t), hipMemcpyDeviceToHost,
                         c.stream));
  OFD_HIP(hipStreamSynchronize(c.stream));
  if (neg > 0) {
    set_error("negative determinant Jacobian at " + std::to_string(neg) +
              " Gauss points (eval_shape_fem warning)");
    return OFD_ERR_NEGDETJ;
  }
  return OFD_OK;
}

int ofd_get_tem_system(ofd_ctx *x, double *lhs, double *rhs) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  Ctx &c = x->c;
  int rc = ensure_tem(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  if (lhs) rc = d2h(c, lhs, c.d_AT, (size_t)c.icnt * sizeof(double));
  if (!rc && rhs) rc = d2h(c, rhs, c.d_rhsT, (size_t)c.nnode * sizeof(double));
  return rc;
}

int ofd_set_tem_system(ofd_ctx *x, const double *lhs, const double *rhs) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  Ctx &c = x->c;
  int rc = ensure_tem(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  if (lhs) {
    rc = h2d(c, c.d_AT, lhs, (size_t)c.icnt * sizeof(double));
    if (rc) return rc;
    c.matT_valid = true;
    c.precT_valid = c.aT32_valid = c.apT32_valid = false;
    ++c.matT_version;
  }
  if (rhs) rc = h2d(c, c.d_rhsT, rhs, (size_t)c.nnode * sizeof(double));
  return rc;
}

int ofd_spmv_tem(ofd_ctx *x, int precision, const double *xin, double *y) {
  OFD_CHECK(x && xin && y, OFD_ERR_ARG, "null argument");
  OFD_CHECK(precision == OFD_FP64 || precision == OFD_FP32, OFD_ERR_ARG, "bad precision");
  Ctx &c = x->c;
  int rc = ensure_tem(c);
  if (rc) return rc;
  OFD_CHECK(c.matT_valid, OFD_ERR_ARG, "no assembled temperature matrix");
  OFD_HIP(hipSetDevice(c.device));
  if (precision == OFD_FP32) {
    rc = ensure_aT32(c);
    if (rc) return rc;
  }
  TemSystem sys(c);
  rc = h2d(c, c.d_yT, xin, (size_t)c.nnode * sizeof(double));
  if (rc) return rc;
  OFD_HIP(launch_spmv_s(c, precision, c.d_yT, c.d_rT, nullptr, 0));
  return d2h(c, y, c.d_rT, (size_t)c.nnode * sizeof(double));
}

int ofd_gmres_tem(ofd_ctx *x, int precision, int ortho, double tol, int kspace,
                  int kspace_min, double *sol, int *iters, double *resid) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  OFD_CHECK(precision == OFD_FP64 || precision == OFD_FP32, OFD_ERR_ARG, "bad precision");
  OFD_CHECK(ortho == OFD_ORTHO_MGS || ortho == OFD_ORTHO_CGS2 || ortho == OFD_ORTHO_CGS1,
            OFD_ERR_ARG, "bad ortho");
  Ctx &c = x->c;
  int rc = ensure_tem(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  TemSystem sys(c);
  OFD_HIP(hipMemsetAsync(c.d_xT, 0, (size_t)c.gm_n * 4 * sizeof(double), c.stream));
  // |b| = 0: the reference returns sol = 0 (SparseGMRES_tem.f90:68-71)
  double bn = 0.0;
  rc = norm2_dev(c, c.d_rhsT, &bn);
  if (rc) return rc;
  OFD_CHECK(std::isfinite(bn), OFD_ERR_NONFINITE, "SparseGMRES_tem: non-finite right-hand side");
  int it = 0;
  double res = 0.0;
  if (bn > 0.0) {
    OFD_HIP(launch_outer_reset(c, 0.0));
    rc = gmres_cycle(c, precision, ortho, c.d_rhsT, tol, kspace, kspace_min, c.d_xT, &it,
                     &res, nullptr);
  }
  if (iters) *iters = it;
  if (resid) *resid = res;
  if (rc) return rc;
  OFD_HIP(hipMemcpyAsync(c.d_field[OFD_SOLT], c.d_xT, (size_t)c.nnode * sizeof(double),
                         hipMemcpyDeviceToDevice, c.stream));
  if (sol) rc = d2h(c, sol, c.d_xT, (size_t)c.nnode * sizeof(double));
  OFD_HIP(hipStreamSynchronize(c.stream));
  return rc;
}

int ofd_solve_tem(ofd_ctx *x, int precision, int ortho, double rtol, double inner_tol,
                  int kspace, int max_cycles, int *iters, int *cycles, double *relres) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  OFD_CHECK(precision == OFD_FP64 || precision == OFD_FP32, OFD_ERR_ARG, "bad precision");
  OFD_CHECK(max_cycles >= 1, OFD_ERR_ARG, "max_cycles >= 1");
  Ctx &c = x->c;
  int rc = ensure_tem(c);
  if (rc) return rc;
  OFD_CHECK(c.matT_valid, OFD_ERR_ARG, "no assembled temperature matrix");
  OFD_HIP(hipSetDevice(c.device));
  TemSystem sys(c);
  rc = solve_core(c, precision, ortho, c.d_rhsT, c.d_xT, c.d_rT, rtol, inner_tol, kspace,
                  max_cycles, iters, cycles, relres);
  if (rc && rc != OFD_ERR_NOCONV) return rc;
  OFD_HIP(hipMemcpyAsync(c.d_field[OFD_SOLT], c.d_xT, (size_t)c.nnode * sizeof(double),
                         hipMemcpyDeviceToDevice, c.stream));
  OFD_HIP(hipStreamSynchronize(c.stream));
  return rc;
}

int ofd_tem_update(ofd_ctx *x) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  Ctx &c = x->c;
  int rc = need_mesh(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  OFD_HIP(launch_fields_s(c, 0, c.d_field[OFD_SOLT], OFD_TG, OFD_TGOLD, OFD_RTG,
                          OFD_RTGOLD));
  OFD_HIP(hipStreamSynchronize(c.stream));
  return OFD_OK;
}

int ofd_tem_rhs_norm(ofd_ctx *x, double *norm2) {
  OFD_CHECK(x && norm2, OFD_ERR_ARG, "null argument");
  Ctx &c = x->c;
  int rc = ensure_tem(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  TemSystem sys(c);
  return norm2_dev(c, c.d_rhsT, norm2);
}

// ---- coupled NS + VOF system ----

static int ensure_nsvof(Ctx &c) {
  int rc = need_pattern(c);
  if (rc) return rc;
  OFD_CHECK((size_t)5 * c.nnode <= (size_t)4 * kMaxRedBlocks * kRedBlock, OFD_ERR_UNSUPPORTED,
            "the NSVOF system supports at most 838,860 nodes per context");
  const int n4 = (5 * c.nnode + 3) / 4;
  if (c.nsvV_icnt != c.icnt) {
    rc = dalloc(c, &c.d_AV, (size_t)c.icnt * 24);
    if (!rc) rc = dalloc(c, &c.d_AV32, (size_t)0);
    if (!rc) rc = dalloc(c, &c.d_ApV, (size_t)0);
    if (!rc) rc = dalloc(c, &c.d_ApV32, (size_t)0);
    if (rc) return rc;
    OFD_HIP(hipMemsetAsync(c.d_AV, 0, (size_t)c.icnt * 24 * sizeof(double), c.stream));
    c.nsvV_icnt = c.icnt;
    c.matV_valid = c.precV_valid = c.pV_valid = c.aV32_valid = c.apV32_valid = false;
  }
  if (c.nsvV_n4 != n4) {
    const size_t np = (size_t)n4 * 4;
    rc = dalloc(c, &c.d_rhsV, np);
    if (!rc) rc = dalloc(c, &c.d_xV, np);
    if (!rc) rc = dalloc(c, &c.d_rV, np);
    if (!rc) rc = dalloc(c, &c.d_yV, np);
    if (!rc) rc = dalloc(c, &c.d_PV, (size_t)c.nnode * 17);
    if (rc) return rc;
    OFD_HIP(hipMemsetAsync(c.d_rhsV, 0, np * sizeof(double), c.stream));
    OFD_HIP(hipMemsetAsync(c.d_xV, 0, np * sizeof(double), c.stream));
    OFD_HIP(hipMemsetAsync(c.d_rV, 0, np * sizeof(double), c.stream));
    OFD_HIP(hipMemsetAsync(c.d_yV, 0, np * sizeof(double), c.stream));
    c.nsvV_n4 = n4;
  }
  return OFD_OK;
}

// reference arrays (LHSK11 (9,Z), LHSG/LHSD1/LHSLSU/LHSULS (3,Z), LHSM/LHSLS/
// LHSPLS (Z)) -> the device layout: Z blocks [K11 G; D1 M] of 16 row-major,
// then Z pair groups of 8, (ULS(r), LSU(r)) r < 3, (PLS, LS)
static void nsvof_to_device(int64_t Z, const double *k11, const double *g, const double *d1,
                            const double *m, const double *ls, const double *lsu,
                            const double *pls, const double *uls, double *A) {
  for (int64_t k = 0; k < Z; ++k) {
    double *b = A + k * 16, *e = A + 16 * Z + k * 8;
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) b[4 * r + c] = k11[k * 9 + r * 3 + c];
      b[4 * r + 3] = g[k * 3 + r];
      b[12 + r] = d1[k * 3 + r];
      e[2 * r] = uls[k * 3 + r];
      e[2 * r + 1] = lsu[k * 3 + r];
    }
    b[15] = m[k];
    e[6] = pls[k];
    e[7] = ls[k];
  }
}

// node vector: reference (u (N,3) column-major, p, phi) <-> [N][4] then [N]
static void nsvof_pack(int N, const double *u, const double *p, const double *phi,
                       double *v) {
  for (int i = 0; i < N; ++i) {
    for (int c = 0; c < 3; ++c) v[(size_t)i * 4 + c] = u[(size_t)c * N + i];
    v[(size_t)i * 4 + 3] = p[i];
    v[(size_t)4 * N + i] = phi[i];
  }
}

static void nsvof_unpack(int N, const double *v, double *u, double *p, double *phi) {
  for (int i = 0; i < N; ++i) {
    if (u)
      for (int c = 0; c < 3; ++c) u[(size_t)c * N + i] = v[(size_t)i * 4 + c];
    if (p) p[i] = v[(size_t)i * 4 + 3];
    if (phi) phi[i] = v[(size_t)4 * N + i];
  }
}

int ofd_assemble_nsvof(ofd_ctx *x, int tensor_flag, double *cfl) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  Ctx &c = x->c;
  int rc = ensure_nsvof(c);
  if (rc) return rc;
  OFD_CHECK(c.quench, OFD_ERR_ARG, "no quench parameters (ofd_set_quench_params)");
  OFD_CHECK(c.nshl == 8 || c.nshl == 4 || c.nshl == 6, OFD_ERR_UNSUPPORTED,
            "the GPU element assembly is implemented for trilinear hexes, linear tets and "
            "prisms");
  OFD_CHECK((tensor_flag & ~(OFD_ASSEMBLE_VEC | OFD_ASSEMBLE_MAT)) == 0, OFD_ERR_ARG,
            "tensor_flag must combine OFD_ASSEMBLE_VEC / OFD_ASSEMBLE_MAT");
  OFD_HIP(hipSetDevice(c.device));
  OFD_HIP(launch_alpha_stage(c, true));
  if (tensor_flag & OFD_ASSEMBLE_VEC)
    OFD_HIP(launch_zero(c, c.d_rhsV, (size_t)c.nsvV_n4 * 4));
  if (tensor_flag & OFD_ASSEMBLE_MAT) {
    // no zeroing: the first element to reach a nonzero stores into it
    c.matV_valid = true;
    c.precV_valid = c.pV_valid = c.aV32_valid = c.apV32_valid = false;
    c.t_matV_valid = false;
  }
  OFD_HIP(hipMemsetAsync(c.d_negdet, 0, sizeof(int32_t), c.stream));
  for (int col = 0; col + 1 < (int)c.color_off.size(); ++col)
    OFD_HIP(launch_assemble_nsvof_hex(c, tensor_flag, col));
  double *d_out = c.d_part;  // scratch for the CFL maximum
  OFD_HIP(launch_max_reduce(c, c.d_cfl, c.nelem, d_out));
  double h[1] = {0.0};
  int32_t neg = 0;
  OFD_HIP(hipMemcpyAsync(h, d_out, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  OFD_HIP(hipMemcpyAsync(&neg, c.d_negdet, sizeof(int32_t), hipMemcpyDeviceToHost,
                         c.stream));
  OFD_HIP(hipStreamSynchronize(c.stream));
  if (cfl) *cfl = h[0];
  if (neg > 0) {
    set_error("negative determinant Jacobian at " + std::to_string(neg) +
              " Gauss points (eval_shape_fem warning)");
    return OFD_ERR_NEGDETJ;
  }
  return OFD_OK;
}

int ofd_get_nsvof_system(ofd_ctx *x, double *k11, double *g, double *d1, double *m,
                         double *ls, double *lsu, double *pls, double *uls, double *rhsu,
                         double *rhsp, double *rhsls) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  Ctx &c = x->c;
  int rc = ensure_nsvof(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  const int64_t Z = c.icnt;
  const int N = c.nnode;
  if (k11 || g || d1 || m || ls || lsu || pls || uls) {
    std::vector<double> A((size_t)Z * 24);
    rc = d2h(c, A.data(), c.d_AV, A.size() * sizeof(double));
    if (rc) return rc;
    for (int64_t k = 0; k < Z; ++k) {
      const double *b = A.data() + k * 16, *e = A.data() + 16 * Z + k * 8;
      for (int r = 0; r < 3; ++r) {
        for (int cc = 0; cc < 3; ++cc)
          if (k11) k11[k * 9 + r * 3 + cc] = b[4 * r + cc];
        if (g) g[k * 3 + r] = b[4 * r + 3];
        if (d1) d1[k * 3 + r] = b[12 + r];
        if (uls) uls[k * 3 + r] = e[2 * r];
        if (lsu) lsu[k * 3 + r] = e[2 * r + 1];
      }
      if (m) m[k] = b[15];
      if (pls) pls[k] = e[6];
      if (ls) ls[k] = e[7];
    }
  }
  if (rhsu || rhsp || rhsls) {
    std::vector<double> r((size_t)N * 5);
    rc = d2h(c, r.data(), c.d_rhsV, r.size() * sizeof(double));
    if (rc) return rc;
    nsvof_unpack(N, r.data(), rhsu, rhsp, rhsls);
  }
  return OFD_OK;
}

int ofd_set_nsvof_system(ofd_ctx *x, const double *k11, const double *g, const double *d1,
                         const double *m, const double *ls, const double *lsu,
                         const double *pls, const double *uls, const double *rhsu,
                         const double *rhsp, const double *rhsls) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  Ctx &c = x->c;
  int rc = ensure_nsvof(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  const int64_t Z = c.icnt;
  const int N = c.nnode;
  if (k11) {
    OFD_CHECK(g && d1 && m && ls && lsu && pls && uls, OFD_ERR_ARG,
              "the matrix needs all eight arrays");
    std::vector<double> A((size_t)Z * 24);
    nsvof_to_device(Z, k11, g, d1, m, ls, lsu, pls, uls, A.data());
    rc = h2d(c, c.d_AV, A.data(), A.size() * sizeof(double));
    if (rc) return rc;
    c.matV_valid = true;
    c.precV_valid = c.pV_valid = c.aV32_valid = c.apV32_valid = false;
    c.t_matV_valid = false;
  }
  if (rhsu) {
    OFD_CHECK(rhsp && rhsls, OFD_ERR_ARG, "the right-hand side needs all three arrays");
    std::vector<double> r((size_t)N * 5);
    nsvof_pack(N, rhsu, rhsp, rhsls, r.data());
    rc = h2d(c, c.d_rhsV, r.data(), r.size() * sizeof(double));
  }
  return rc;
}

int ofd_spmv_nsvof(ofd_ctx *x, int precision, const double *xu, const double *xp,
                   const double *xl, double *yu, double *yp, double *yl) {
  OFD_CHECK(x && xu && xp && xl && yu && yp && yl, OFD_ERR_ARG, "null argument");
  OFD_CHECK(precision == OFD_FP64 || precision == OFD_FP32, OFD_ERR_ARG, "bad precision");
  Ctx &c = x->c;
  int rc = ensure_nsvof(c);
  if (rc) return rc;
  OFD_CHECK(c.matV_valid, OFD_ERR_ARG, "no assembled NSVOF matrix");
  OFD_HIP(hipSetDevice(c.device));
  if (precision == OFD_FP32) {
    rc = ensure_aV32(c);
    if (rc) return rc;
  }
  const int N = c.nnode;
  std::vector<double> v((size_t)N * 5);
  nsvof_pack(N, xu, xp, xl, v.data());
  NsvofSystem sys(c);
  rc = h2d(c, c.d_yV, v.data(), v.size() * sizeof(double));
  if (rc) return rc;
  OFD_HIP(launch_spmv_v(c, precision, c.d_yV, c.d_rV, nullptr, 0));
  rc = d2h(c, v.data(), c.d_rV, v.size() * sizeof(double));
  if (rc) return rc;
  nsvof_unpack(N, v.data(), yu, yp, yl);
  return OFD_OK;
}

int ofd_gmres_nsvof(ofd_ctx *x, int precision, int ortho, double tol, int kspace,
                    int kspace_min, double *solu, double *solp, double *solls, int *iters,
                    double *resid) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  OFD_CHECK(precision == OFD_FP64 || precision == OFD_FP32, OFD_ERR_ARG, "bad precision");
  OFD_CHECK(ortho == OFD_ORTHO_MGS || ortho == OFD_ORTHO_CGS2 || ortho == OFD_ORTHO_CGS1,
            OFD_ERR_ARG, "bad ortho");
  Ctx &c = x->c;
  int rc = ensure_nsvof(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  NsvofSystem sys(c);
  OFD_HIP(hipMemsetAsync(c.d_xV, 0, (size_t)c.gm_n * 4 * sizeof(double), c.stream));
  double bn = 0.0;
  rc = norm2_dev(c, c.d_rhsV, &bn);
  if (rc) return rc;
  OFD_CHECK(std::isfinite(bn), OFD_ERR_NONFINITE,
            "SparseGMRES_ns_conv: non-finite right-hand side");
  OFD_HIP(launch_outer_reset(c, 0.0));
  int it = 0;
  double res = 0.0;
  rc = gmres_cycle(c, precision, ortho, c.d_rhsV, tol, kspace, kspace_min, c.d_xV, &it, &res,
                   nullptr);
  if (iters) *iters = it;
  if (resid) *resid = res;
  if (rc) return rc;
  c.sol_src = 2;
  OFD_HIP(launch_nsvof_to_soa(c, c.d_xV));
  const size_t nb = (size_t)c.nnode * sizeof(double);
  if (solu) rc = d2h(c, solu, c.d_field[OFD_SOLU], 3 * nb);
  if (!rc && solp) rc = d2h(c, solp, c.d_field[OFD_SOLP], nb);
  if (!rc && solls) rc = d2h(c, solls, c.d_field[OFD_SOLPHI], nb);
  OFD_HIP(hipStreamSynchronize(c.stream));
  return rc;
}

int ofd_solve_nsvof(ofd_ctx *x, int precision, int ortho, double rtol, double inner_tol,
                    int kspace, int max_cycles, int *iters, int *cycles, double *relres) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  OFD_CHECK(precision == OFD_FP64 || precision == OFD_FP32, OFD_ERR_ARG, "bad precision");
  OFD_CHECK(max_cycles >= 1, OFD_ERR_ARG, "max_cycles >= 1");
  Ctx &c = x->c;
  int rc = ensure_nsvof(c);
  if (rc) return rc;
  OFD_CHECK(c.matV_valid, OFD_ERR_ARG, "no assembled NSVOF matrix");
  OFD_HIP(hipSetDevice(c.device));
  NsvofSystem sys(c);
  rc = solve_core(c, precision, ortho, c.d_rhsV, c.d_xV, c.d_rV, rtol, inner_tol, kspace,
                  max_cycles, iters, cycles, relres);
  if (rc && rc != OFD_ERR_NOCONV) return rc;
  c.sol_src = 2;
  OFD_HIP(launch_nsvof_to_soa(c, c.d_xV));
  OFD_HIP(hipStreamSynchronize(c.stream));
  return rc;
}

int ofd_nsvof_update(ofd_ctx *x) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  Ctx &c = x->c;
  int rc = ensure_nsvof(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  OFD_HIP(launch_fields_v(c));
  OFD_HIP(hipStreamSynchronize(c.stream));
  return OFD_OK;
}

int ofd_nsvof_rhs_norms(ofd_ctx *x, double *norms3) {
  OFD_CHECK(x && norms3, OFD_ERR_ARG, "null argument");
  Ctx &c = x->c;
  int rc = ensure_nsvof(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  OFD_HIP(launch_norm2_v(c, c.d_rhsV));
  OFD_HIP(hipMemcpyAsync(c.h_gs, c.d_gs, sizeof(GmresState), hipMemcpyDeviceToHost,
                         c.stream));
  OFD_HIP(hipStreamSynchronize(c.stream));
  for (int j = 0; j < 3; ++j) norms3[j] = c.h_gs->scal[1 + j];
  return OFD_OK;
}

static int ensure_fringe(Ctx &c, int nrecv) {
  int rc = OFD_OK;
  if (c.fr_alloc_e < c.nelem) {
    rc = dalloc(c, &c.d_fr_box, (size_t)c.nelem * 6);
    if (!rc) rc = dalloc(c, &c.d_fr_ok, (size_t)c.nelem);
    if (!rc) rc = dalloc(c, &c.d_fr_geo, (size_t)c.nelem * 24);
    if (rc) return rc;
    c.fr_alloc_e = c.nelem;
  }
  if (c.fr_alloc_r < nrecv) {
    rc = dalloc(c, &c.d_fr_xr, (size_t)nrecv * 3);
    if (!rc) rc = dalloc(c, &c.d_fr_dr, (size_t)nrecv * 3);
    if (!rc) rc = dalloc(c, &c.d_fr_val, (size_t)nrecv * 3);
    if (!rc) rc = dalloc(c, &c.d_fr_donor, (size_t)nrecv);
    if (!rc) rc = dalloc(c, &c.d_fr_nodes, (size_t)nrecv);
    if (rc) return rc;
    c.fr_alloc_r = nrecv;
  }
  return OFD_OK;
}

int ofd_fringe_interp(ofd_ctx *x, int32_t src_id, int nrecv, const double *xr,
                      const double *dr, int32_t *donor, double *val) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  OFD_CHECK(nrecv >= 0, OFD_ERR_ARG, "nrecv < 0");
  Ctx &c = x->c;
  int rc = need_mesh(c);
  if (rc) return rc;
  if (nrecv == 0) return OFD_OK;
  OFD_CHECK(xr && dr && donor && val, OFD_ERR_ARG, "null argument");
  OFD_CHECK(c.nshl == 4 || c.nshl == 8, OFD_ERR_UNSUPPORTED,
            "the donor search is implemented for linear tets and trilinear hexes");
  OFD_HIP(hipSetDevice(c.device));
  rc = ensure_fringe(c, nrecv);
  if (rc) return rc;
  const size_t vb = (size_t)nrecv * 3 * sizeof(double);
  rc = h2d(c, c.d_fr_xr, xr, vb);
  if (!rc) rc = h2d(c, c.d_fr_dr, dr, vb);
  if (rc) return rc;
  OFD_HIP(launch_fringe(c, src_id, nrecv, c.d_fr_xr, c.d_fr_dr, c.d_fr_donor, c.d_fr_val,
                        c.d_fr_box, c.d_fr_ok, c.d_fr_geo));
  rc = d2h(c, donor, c.d_fr_donor, (size_t)nrecv * sizeof(int32_t));
  if (!rc) rc = d2h(c, val, c.d_fr_val, vb);
  if (rc) return rc;
  OFD_HIP(hipStreamSynchronize(c.stream));
  for (int j = 0; j < nrecv; ++j) donor[j] += 1;  // 1-based, 0 = no donor
  return OFD_OK;
}

int ofd_apply_fringe(ofd_ctx *x, int nrecv, const int32_t *nodes, const double *val) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  OFD_CHECK(nrecv >= 0, OFD_ERR_ARG, "nrecv < 0");
  Ctx &c = x->c;
  int rc = need_mesh(c);
  if (rc) return rc;
  if (nrecv == 0) return OFD_OK;
  OFD_CHECK(nodes && val, OFD_ERR_ARG, "null argument");
  // distinct nodes in range: one thread writes each node
  std::vector<int32_t> nd(nodes, nodes + nrecv);
  std::vector<uint8_t> seen((size_t)c.nnode, 0);
  for (int j = 0; j < nrecv; ++j) {
    const int k = nd[j] - 1;
    OFD_CHECK(k >= 0 && k < c.nnode, OFD_ERR_ARG, "fringe node out of range [1, NNODE]");
    OFD_CHECK(!seen[k], OFD_ERR_ARG, "fringe nodes must be distinct");
    seen[k] = 1;
    nd[j] = k;
  }
  OFD_HIP(hipSetDevice(c.device));
  rc = ensure_fringe(c, nrecv);
  if (rc) return rc;
  rc = h2d(c, c.d_fr_nodes, nd.data(), (size_t)nrecv * sizeof(int32_t));
  if (!rc) rc = h2d(c, c.d_fr_val, val, (size_t)nrecv * 3 * sizeof(double));
  if (rc) return rc;
  OFD_HIP(launch_fringe_apply(c, nrecv, c.d_fr_nodes, c.d_fr_val));
  OFD_HIP(hipStreamSynchronize(c.stream));
  return OFD_OK;
}

int ofd_time_copy(ofd_ctx *x, int64_t nbytes, int reps, double *ms) {
  OFD_CHECK(x && ms && reps > 0, OFD_ERR_ARG, "bad argument");
  OFD_CHECK(nbytes >= 16 && nbytes % 16 == 0 && nbytes <= (int64_t(1) << 34), OFD_ERR_ARG,
            "nbytes must be a multiple of 16 in [16, 16 GiB]");
  Ctx &c = x->c;
  OFD_HIP(hipSetDevice(c.device));
  void *src = nullptr, *dst = nullptr;
  OFD_HIP(hipMalloc(&src, (size_t)nbytes));
  hipError_t e = hipMalloc(&dst, (size_t)nbytes);
  if (e != hipSuccess) {
    (void)hipFree(src);
    OFD_HIP(e);
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float t = 0.f;
  e = hipMemsetAsync(src, 0, (size_t)nbytes, c.stream);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = launch_copy(c, src, dst, (size_t)nbytes);  // warm
  if (e == hipSuccess) e = hipEventRecord(e0, c.stream);
  for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_copy(c, src, dst, (size_t)nbytes);
  if (e == hipSuccess) e = hipEventRecord(e1, c.stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(src);
  (void)hipFree(dst);
  OFD_HIP(e);
  *ms = (double)t / reps;
  return OFD_OK;
}

int ofd_time_read(ofd_ctx *x, int64_t nbytes, int reps, double *ms) {
  OFD_CHECK(x && ms && reps > 0, OFD_ERR_ARG, "bad argument");
  OFD_CHECK(nbytes >= 16 && nbytes % 16 == 0 && nbytes <= (int64_t(1) << 35), OFD_ERR_ARG,
            "nbytes must be a multiple of 16 in [16, 32 GiB]");
  Ctx &c = x->c;
  OFD_HIP(hipSetDevice(c.device));
  void *src = nullptr;
  double *out = nullptr;
  OFD_HIP(hipMalloc(&src, (size_t)nbytes));
  hipError_t e = hipMalloc(reinterpret_cast<void **>(&out), sizeof(double) * kReadBlocks);
  if (e != hipSuccess) {
    (void)hipFree(src);
    OFD_HIP(e);
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float t = 0.f;
  e = hipMemsetAsync(src, 0, (size_t)nbytes, c.stream);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = launch_read(c, src, (size_t)nbytes, out);  // warm
  if (e == hipSuccess) e = hipEventRecord(e0, c.stream);
  for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_read(c, src, (size_t)nbytes, out);
  if (e == hipSuccess) e = hipEventRecord(e1, c.stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(src);
  (void)hipFree(out);
  OFD_HIP(e);
  *ms = (double)t / reps;
  return OFD_OK;
}

int ofd_probe_set(ofd_ctx *x, int every) {
  OFD_CHECK(x && every >= 0, OFD_ERR_ARG, "bad argument");
  Ctx &c = x->c;
  OFD_HIP(hipSetDevice(c.device));
  if (every > 0 && c.probe_ev.empty()) {
    // read back when a solve returns; launches past the capacity of one
    // solve are not sampled
    c.probe_ev.resize(2 * kProbeSlots);
    for (auto &ev : c.probe_ev) OFD_HIP(hipEventCreate(&ev));
  }
  c.probe_every = every;
  c.probe_used = 0;
  c.probe_tick = 0;
  return OFD_OK;
}

int ofd_probe_read(ofd_ctx *x, int64_t *count, double *total_ms) {
  OFD_CHECK(x && count && total_ms, OFD_ERR_ARG, "null argument");
  Ctx &c = x->c;
  *count = c.probe_count;
  *total_ms = c.probe_ms;
  c.probe_count = 0;
  c.probe_ms = 0.0;
  return OFD_OK;
}

// ---- adjoint of the NS Newton step ----

// adjoint vectors and scratch, allocated on first use
static int ensure_adjoint(Ctx &c) {
  int rc = need_pattern(c);
  if (rc) return rc;
  OFD_CHECK(c.h_face_e.empty(), OFD_ERR_UNSUPPORTED,
            "the adjoint differentiates the volume residual; this context has weak "
            "boundary faces (ofd_set_weak_faces)");
  const size_t n4 = (size_t)c.nnode * 4;
  double **bufs[5] = {&c.d_jt, &c.d_lam, &c.d_lamr, &c.d_lamb, &c.d_lamd};
  for (double **b : bufs)
    if (!*b) {
      rc = dalloc(c, b, n4);
      if (rc) return rc;
      OFD_HIP(hipMemsetAsync(*b, 0, n4 * sizeof(double), c.stream));
    }
  if (!c.d_vjp) {
    rc = dalloc(c, &c.d_vjp, (size_t)c.nnode * 8);
    if (rc) return rc;
  }
  return OFD_OK;
}

// t_A64 = the assembled LHS transposed (blocks mirrored through d_tmap)
static int ensure_transpose(Ctx &c) {
  OFD_CHECK(c.mat_valid, OFD_ERR_ARG, "no assembled matrix");
  if (c.t_mat_valid && c.t_version == c.mat_version) return OFD_OK;
  int rc = OFD_OK;
  if (!c.t_A64) rc = dalloc(c, &c.t_A64, (size_t)c.icnt * 16);
  if (!rc && !c.t_P64) rc = dalloc(c, &c.t_P64, (size_t)c.nnode * 16);
  if (rc) return rc;
  OFD_HIP(launch_block_transpose(c));
  c.t_mat_valid = true;
  c.t_prec_valid = c.t_a32_valid = c.t_ap32_valid = false;
  c.t_version = c.mat_version;
  return OFD_OK;
}

// the stored-matrix kernels see LHS^T (and its own FP32, block-Jacobi and
// preconditioned copies) while a view is alive
struct TransposeView {
  Ctx &c;
  explicit TransposeView(Ctx &cc) : c(cc) { swap_all(); }
  ~TransposeView() { swap_all(); }
  void swap_all() {
    std::swap(c.d_A64, c.t_A64);
    std::swap(c.d_A32, c.t_A32);
    std::swap(c.d_P64, c.t_P64);
    std::swap(c.d_Ap64, c.t_Ap64);
    std::swap(c.d_Ap32, c.t_Ap32);
    std::swap(c.mat_valid, c.t_mat_valid);
    std::swap(c.prec_valid, c.t_prec_valid);
    std::swap(c.a32_valid, c.t_a32_valid);
    std::swap(c.ap32_valid, c.t_ap32_valid);
  }
};

static int put_node4(Ctx &c, const double *u, const double *p, double *dst) {
  double *su = c.d_field[OFD_LAMU], *sp = c.d_field[OFD_LAMP];
  int rc = h2d(c, su, u, (size_t)c.nnode * 3 * sizeof(double));
  if (!rc) rc = h2d(c, sp, p, (size_t)c.nnode * sizeof(double));
  if (rc) return rc;
  OFD_HIP(launch_soa_to_node4(c, su, sp, dst));
  return OFD_OK;
}

static int get_node4(Ctx &c, const double *src, double *u, double *p) {
  double *su = c.d_field[OFD_LAMU], *sp = c.d_field[OFD_LAMP];
  OFD_HIP(launch_node4_to_soa(c, src, su, sp));
  int rc = d2h(c, u, su, (size_t)c.nnode * 3 * sizeof(double));
  if (!rc) rc = d2h(c, p, sp, (size_t)c.nnode * sizeof(double));
  return rc;
}

int ofd_spmv_up_t(ofd_ctx *x, int precision, const double *xu, const double *xp,
                  double *yu, double *yp) {
  OFD_CHECK(x && xu && xp && yu && yp, OFD_ERR_ARG, "null argument");
  Ctx &c = x->c;
  int rc = need_pattern(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  rc = ensure_transpose(c);
  if (rc) return rc;
  TransposeView view(c);
  return ofd_spmv_up(x, precision, xu, xp, yu, yp);
}

int ofd_solve_t(ofd_ctx *x, int precision, int ortho, double rtol, double inner_tol,
                int kspace, int max_cycles, const double *gu, const double *gp, int *iters,
                int *cycles, double *relres) {
  OFD_CHECK(x && gu && gp, OFD_ERR_ARG, "null argument");
  OFD_CHECK(precision == OFD_FP64 || precision == OFD_FP32, OFD_ERR_ARG, "bad precision");
  OFD_CHECK(!(ortho & OFD_PREC_RIGHT), OFD_ERR_ARG,
            "OFD_PREC_RIGHT: ofd_solve and ofd_solve_nsvof only");
  OFD_CHECK(max_cycles >= 1, OFD_ERR_ARG, "max_cycles >= 1");
  Ctx &c = x->c;
  int rc = ensure_adjoint(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  rc = ensure_transpose(c);
  if (rc) return rc;
  rc = put_node4(c, gu, gp, c.d_lamb);
  if (rc) return rc;
  {
    TransposeView view(c);
    rc = solve_core(c, precision, ortho, c.d_lamb, c.d_lam, c.d_lamr, rtol, inner_tol, kspace,
                    max_cycles, iters, cycles, relres);
  }
  if (rc && rc != OFD_ERR_NOCONV) return rc;
  OFD_HIP(launch_node4_to_soa(c, c.d_lam, c.d_field[OFD_LAMU], c.d_field[OFD_LAMP]));
  OFD_HIP(hipStreamSynchronize(c.stream));
  return rc;
}

int ofd_vjp_ns(ofd_ctx *x, int subdomain_flag, const double *lu, const double *lp,
               double *gu, double *gac, double *gp, double *gmus) {
  OFD_CHECK(x && lu && lp, OFD_ERR_ARG, "null argument");
  OFD_CHECK(subdomain_flag >= 1 && subdomain_flag <= 3, OFD_ERR_ARG, "subdomain_flag 1..3");
  Ctx &c = x->c;
  int rc = ensure_adjoint(c);
  if (rc) return rc;
  OFD_CHECK(c.nshl == 8 || c.nshl == 4 || c.nshl == 6, OFD_ERR_UNSUPPORTED,
            "the NS adjoint kernels are implemented for trilinear hexes, linear tets and prisms");
  OFD_HIP(hipSetDevice(c.device));
  rc = put_node4(c, lu, lp, c.d_lamb);
  if (rc) return rc;
  OFD_HIP(launch_alpha_stage(c));
  OFD_HIP(launch_zero(c, c.d_vjp, (size_t)c.nnode * 8));
  OFD_HIP(launch_vjp_ns(c, subdomain_flag, c.d_lamb));
  const size_t N = c.nnode;
  if (gu) rc = d2h(c, gu, c.d_vjp, 3 * N * sizeof(double));
  if (!rc && gac) rc = d2h(c, gac, c.d_vjp + 3 * N, 3 * N * sizeof(double));
  if (!rc && gp) rc = d2h(c, gp, c.d_vjp + 6 * N, N * sizeof(double));
  if (!rc && gmus) rc = d2h(c, gmus, c.d_vjp + 7 * N, N * sizeof(double));
  return rc;
}

int ofd_jt_apply(ofd_ctx *x, int subdomain_flag, const double *lu, const double *lp,
                 double *yu, double *yp) {
  OFD_CHECK(x && lu && lp && yu && yp, OFD_ERR_ARG, "null argument");
  OFD_CHECK(subdomain_flag >= 1 && subdomain_flag <= 3, OFD_ERR_ARG, "subdomain_flag 1..3");
  Ctx &c = x->c;
  int rc = ensure_adjoint(c);
  if (rc) return rc;
  OFD_CHECK(c.nshl == 8 || c.nshl == 4 || c.nshl == 6, OFD_ERR_UNSUPPORTED,
            "the NS adjoint kernels are implemented for trilinear hexes, linear tets and prisms");
  OFD_HIP(hipSetDevice(c.device));
  rc = put_node4(c, lu, lp, c.d_lamb);
  if (rc) return rc;
  OFD_HIP(launch_alpha_stage(c));
  OFD_HIP(launch_jt_ns(c, subdomain_flag, c.d_lamb, c.d_jt));
  rc = get_node4(c, c.d_jt, yu, yp);
  OFD_HIP(hipStreamSynchronize(c.stream));
  return rc;
}

// Restarted flexible GMRES (right preconditioning) on vectors of n doubles
// (the current system's gm_n 4-tuples): apply(v, y) is the operator,
// prec(v, z, &inner_iters) an inner solve z ~ M^-1 v.  The outer basis is
// FP64 with classical Gram-Schmidt applied twice; x accumulates; w is an
// n-vector of scratch.
extern "C++" {
template <class Apply, class Prec>
static int fgmres(Ctx &c, size_t n, Apply &&apply, Prec &&prec, const double *b, double *x,
                  double *w, double rtol, int m, int max_cycles, int *iters, int *cycles,
                  double *relres) {
  OFD_CHECK(m >= 1 && m <= 100, OFD_ERR_ARG, "kspace out of range [1, 100]");
  if (c.fg_m < m || c.fg_n < n) {
    const int mm = std::max(c.fg_m, m);
    const size_t nn = std::max(c.fg_n, n);
    int rc = dalloc(c, &c.d_fv, nn * (mm + 1));
    if (!rc) rc = dalloc(c, &c.d_fz, nn * mm);
    if (!rc) rc = dalloc(c, &c.d_fh, (size_t)4 * (mm + 2));
    if (rc) return rc;
    c.fg_m = mm;
    c.fg_n = nn;
  }
  double *V = c.d_fv, *Z = c.d_fz, *h1 = c.d_fh, *h2 = c.d_fh + (m + 2),
         *nr = c.d_fh + 2 * (m + 2), *yd = c.d_fh + 3 * (m + 2);
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), e(m + 1), y(m), hbuf(2 * m + 3);
  double bnorm = 0.0;
  int rc = norm2_dev(c, b, &bnorm);
  if (rc) return rc;
  OFD_CHECK(std::isfinite(bnorm), OFD_ERR_NONFINITE, "fgmres: non-finite right-hand side");
  int total = 0, cyc = 0;
  double rel = bnorm > 0.0 ? 1.0 : 0.0;
  std::vector<double> hist;
  while (bnorm > 0.0) {
    // true residual r = b - Op x in V_0
    OFD_HIP(apply(x, w));
    OFD_HIP(launch_axpby(c, b, w, -1.0, V, n));
    double beta = 0.0;
    rc = norm2_dev(c, V, &beta);
    if (rc) return rc;
    rel = beta / bnorm;
    hist.push_back(rel);
    if (!std::isfinite(rel)) {
      set_error("fgmres: non-finite residual");
      return OFD_ERR_NONFINITE;
    }
    if (rel <= rtol || cyc >= max_cycles) break;
    ++cyc;
    OFD_HIP(launch_scale(c, n, V, 1.0 / beta, V));
    std::fill(e.begin(), e.end(), 0.0);
    e[0] = beta;
    int k = 0;
    for (; k < m; ++k) {
      double *vk = V + n * k, *zk = Z + n * k;
      int it = 0;
      rc = prec(vk, zk, &it);
      if (rc && rc != OFD_ERR_NOCONV) return rc;
      total += it;
      OFD_HIP(apply(zk, w));
      OFD_HIP(launch_multidot(c, n, w, V, n, k + 1, h1));
      OFD_HIP(launch_multi_axpy(c, n, w, V, n, h1, k + 1, -1.0));
      OFD_HIP(launch_multidot(c, n, w, V, n, k + 1, h2));
      OFD_HIP(launch_multi_axpy(c, n, w, V, n, h2, k + 1, -1.0));
      OFD_HIP(launch_multidot(c, n, w, w, 0, 1, nr));
      OFD_HIP(hipMemcpyAsync(hbuf.data(), h1, sizeof(double) * (k + 1),
                             hipMemcpyDeviceToHost, c.stream));
      OFD_HIP(hipMemcpyAsync(hbuf.data() + m + 1, h2, sizeof(double) * (k + 1),
                             hipMemcpyDeviceToHost, c.stream));
      double hn = 0.0;
      OFD_HIP(hipMemcpyAsync(&hn, nr, sizeof(double), hipMemcpyDeviceToHost, c.stream));
      OFD_HIP(hipStreamSynchronize(c.stream));
      hn = sqrt(hn);
      double *hc = H.data() + (size_t)(m + 1) * k;  // column k
      for (int i = 0; i <= k; ++i) hc[i] = hbuf[i] + hbuf[m + 1 + i];
      hc[k + 1] = hn;
      if (hn > 0.0) OFD_HIP(launch_scale(c, n, w, 1.0 / hn, V + n * (k + 1)));
      for (int i = 0; i < k; ++i) {
        const double t = cs[i] * hc[i] + sn[i] * hc[i + 1];
        hc[i + 1] = -sn[i] * hc[i] + cs[i] * hc[i + 1];
        hc[i] = t;
      }
      const double r = sqrt(hc[k] * hc[k] + hc[k + 1] * hc[k + 1]);
      cs[k] = hc[k] / r;
      sn[k] = hc[k + 1] / r;
      hc[k] = r;
      hc[k + 1] = 0.0;
      e[k + 1] = -sn[k] * e[k];
      e[k] = cs[k] * e[k];
      if (fabs(e[k + 1]) <= rtol * bnorm || hn == 0.0) {
        ++k;
        break;
      }
    }
    // x += Z y, H y = e (upper triangular, k columns)
    for (int j = k - 1; j >= 0; --j) {
      double t = e[j];
      for (int l = j + 1; l < k; ++l) t -= H[(size_t)(m + 1) * l + j] * y[l];
      y[j] = t / H[(size_t)(m + 1) * j + j];
    }
    OFD_HIP(hipMemcpyAsync(yd, y.data(), sizeof(double) * k, hipMemcpyHostToDevice,
                           c.stream));
    OFD_HIP(launch_multi_axpy(c, n, x, Z, n, yd, k, 1.0));
  }
  c.solve_hist = hist;
  OFD_HIP(hipStreamSynchronize(c.stream));
  if (iters) *iters = total;
  if (cycles) *cycles = cyc;
  if (relres) *relres = rel;
  if (rel > rtol) {
    set_error("fgmres: tolerance not reached within max_cycles");
    return OFD_ERR_NOCONV;
  }
  return OFD_OK;
}
}  // extern "C++"

// flexible GMRES on y = J_bc v (op 0) or y = J_bc^T v (op 1), both applied
// exactly by the element kernels, preconditioned by the assembled LHS (op 0)
// or LHS^T (op 1) solved to inner_tol by the restarted device GMRES
static int fgmres_ns(Ctx &c, int op, int sub, int precision, int ortho, const double *b,
                     double *x, double rtol, double inner_tol, int m, int max_cycles,
                     int *iters, int *cycles, double *relres) {
  auto apply = [&](const double *v, double *y) -> hipError_t {
    return op == 0 ? launch_jv_ns(c, sub, v, y) : launch_jt_ns(c, sub, v, y);
  };
  auto prec = [&](const double *v, double *z, int *it) -> int {
    int cy = 0;
    double ir = 0.0;
    const double ctol = precision == OFD_FP32 ? 1e-4 : 1e-2;
    if (op == 1) {
      TransposeView view(c);
      return solve_core(c, precision, ortho, v, z, c.d_lamd, inner_tol, ctol, 30, 100, it,
                        &cy, &ir);
    }
    return solve_core(c, precision, ortho, v, z, c.d_lamd, inner_tol, ctol, 30, 100, it, &cy,
                      &ir);
  };
  return fgmres(c, (size_t)c.nnode * 4, apply, prec, b, x, c.d_jt, rtol, m, max_cycles, iters,
                cycles, relres);
}

int ofd_solve_adjoint(ofd_ctx *x, int subdomain_flag, int precision, int ortho, double rtol,
                      double inner_tol, int kspace, int max_cycles, const double *gu,
                      const double *gp, int *iters, int *cycles, double *relres) {
  OFD_CHECK(x && gu && gp, OFD_ERR_ARG, "null argument");
  OFD_CHECK(subdomain_flag >= 1 && subdomain_flag <= 3, OFD_ERR_ARG, "subdomain_flag 1..3");
  OFD_CHECK(!(ortho & OFD_PREC_RIGHT), OFD_ERR_ARG,
            "OFD_PREC_RIGHT: ofd_solve and ofd_solve_nsvof only");
  OFD_CHECK(precision == OFD_FP64 || precision == OFD_FP32, OFD_ERR_ARG, "bad precision");
  OFD_CHECK(max_cycles >= 1, OFD_ERR_ARG, "max_cycles >= 1");
  Ctx &c = x->c;
  int rc = ensure_adjoint(c);
  if (rc) return rc;
  OFD_CHECK(c.nshl == 8 || c.nshl == 4 || c.nshl == 6, OFD_ERR_UNSUPPORTED,
            "the NS adjoint kernels are implemented for trilinear hexes, linear tets and prisms");
  OFD_HIP(hipSetDevice(c.device));
  rc = ensure_transpose(c);
  if (rc) return rc;
  rc = put_node4(c, gu, gp, c.d_lamb);
  if (rc) return rc;
  OFD_HIP(launch_alpha_stage(c));
  OFD_HIP(launch_zero(c, c.d_lam, (size_t)c.nnode * 4));
  rc = fgmres_ns(c, 1, subdomain_flag, precision, ortho, c.d_lamb, c.d_lam, rtol, inner_tol,
              kspace, max_cycles, iters, cycles, relres);
  if (rc && rc != OFD_ERR_NOCONV) return rc;
  OFD_HIP(launch_node4_to_soa(c, c.d_lam, c.d_field[OFD_LAMU], c.d_field[OFD_LAMP]));
  OFD_HIP(hipStreamSynchronize(c.stream));
  return rc;
}

int ofd_solve_newton(ofd_ctx *x, int subdomain_flag, int precision, int ortho, double rtol,
                     double inner_tol, int kspace, int max_cycles, int *iters, int *cycles,
                     double *relres) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  OFD_CHECK(subdomain_flag >= 1 && subdomain_flag <= 3, OFD_ERR_ARG, "subdomain_flag 1..3");
  OFD_CHECK(!(ortho & OFD_PREC_RIGHT), OFD_ERR_ARG,
            "OFD_PREC_RIGHT: ofd_solve and ofd_solve_nsvof only");
  OFD_CHECK(precision == OFD_FP64 || precision == OFD_FP32, OFD_ERR_ARG, "bad precision");
  OFD_CHECK(max_cycles >= 1, OFD_ERR_ARG, "max_cycles >= 1");
  Ctx &c = x->c;
  int rc = ensure_adjoint(c);
  if (rc) return rc;
  OFD_CHECK(c.mat_valid, OFD_ERR_ARG, "no assembled matrix");
  OFD_CHECK(c.nshl == 8 || c.nshl == 4 || c.nshl == 6, OFD_ERR_UNSUPPORTED,
            "the NS adjoint kernels are implemented for trilinear hexes, linear tets and prisms");
  OFD_HIP(hipSetDevice(c.device));
  OFD_HIP(launch_alpha_stage(c));
  OFD_HIP(launch_zero(c, c.d_x, (size_t)c.nnode * 4));
  rc = fgmres_ns(c, 0, subdomain_flag, precision, ortho, c.d_rhs, c.d_x, rtol, inner_tol, kspace,
              max_cycles, iters, cycles, relres);
  if (rc && rc != OFD_ERR_NOCONV) return rc;
  c.sol_src = 0;
  OFD_HIP(launch_node4_to_soa(c, c.d_x, c.d_field[OFD_SOLU], c.d_field[OFD_SOLP]));
  OFD_HIP(hipStreamSynchronize(c.stream));
  return rc;
}

// ---- temperature adjoint ----

// the adjoint vectors of the T system (zero padded to whole 4-tuples)
static int ensure_tem_adjoint(Ctx &c) {
  int rc = ensure_tem(c);
  if (rc) return rc;
  if (c.tadj_n4 == c.temT_n4 && c.d_vjpT) return OFD_OK;
  free_tem_adjoint(c);
  const size_t np = (size_t)c.temT_n4 * 4;
  double **bufs[4] = {&c.d_lamT, &c.d_lamTb, &c.d_lamTd, &c.d_jtT};
  for (double **b : bufs) {
    rc = dalloc(c, b, np);
    if (rc) return rc;
    OFD_HIP(hipMemsetAsync(*b, 0, np * sizeof(double), c.stream));
  }
  rc = dalloc(c, &c.d_vjpT, (size_t)c.nnode * 7);
  if (rc) return rc;
  c.tadj_n4 = c.temT_n4;
  return OFD_OK;
}

// tT_A = LHSTem transposed (mirrored through d_tmap)
static int ensure_tem_transpose(Ctx &c) {
  OFD_CHECK(c.matT_valid, OFD_ERR_ARG, "no assembled temperature matrix");
  if (c.tT_mat_valid && c.tT_version == c.matT_version) return OFD_OK;
  int rc = OFD_OK;
  if (!c.tT_A) rc = dalloc(c, &c.tT_A, (size_t)c.icnt);
  if (!rc && !c.tT_P) rc = dalloc(c, &c.tT_P, (size_t)c.nnode);
  if (rc) return rc;
  OFD_HIP(launch_scalar_transpose(c));
  c.tT_mat_valid = true;
  c.tT_prec_valid = c.tT_a32_valid = c.tT_ap32_valid = false;
  c.tT_version = c.matT_version;
  return OFD_OK;
}

// the T-system kernels see LHSTem^T (and its own copies) while a view lives
struct TemTransposeView {
  Ctx &c;
  explicit TemTransposeView(Ctx &cc) : c(cc) { swap_all(); }
  ~TemTransposeView() { swap_all(); }
  void swap_all() {
    std::swap(c.d_AT, c.tT_A);
    std::swap(c.d_AT32, c.tT_A32);
    std::swap(c.d_PT, c.tT_P);
    std::swap(c.d_ApT, c.tT_Ap);
    std::swap(c.d_ApT32, c.tT_Ap32);
    std::swap(c.matT_valid, c.tT_mat_valid);
    std::swap(c.precT_valid, c.tT_prec_valid);
    std::swap(c.aT32_valid, c.tT_a32_valid);
    std::swap(c.apT32_valid, c.tT_ap32_valid);
  }
};

static int tem_adjoint_ready(Ctx &c) {
  OFD_HIP(hipSetDevice(c.device));
  int rc = ensure_tem_adjoint(c);
  if (rc) return rc;
  OFD_CHECK(c.quench, OFD_ERR_ARG, "no quench parameters (ofd_set_quench_params)");
  OFD_CHECK(c.nshl == 8 || c.nshl == 4 || c.nshl == 6, OFD_ERR_UNSUPPORTED,
            "the temperature adjoint is implemented for trilinear hexes, linear tets and prisms");
  return OFD_OK;
}

int ofd_spmv_tem_t(ofd_ctx *x, int precision, const double *xin, double *y) {
  OFD_CHECK(x && xin && y, OFD_ERR_ARG, "null argument");
  Ctx &c = x->c;
  int rc = ensure_tem(c);
  if (rc) return rc;
  OFD_HIP(hipSetDevice(c.device));
  rc = ensure_tem_transpose(c);
  if (rc) return rc;
  TemTransposeView view(c);
  return ofd_spmv_tem(x, precision, xin, y);
}

int ofd_solve_tem_t(ofd_ctx *x, int precision, int ortho, double rtol, double inner_tol,
                    int kspace, int max_cycles, const double *g, int *iters, int *cycles,
                    double *relres) {
  OFD_CHECK(x && g, OFD_ERR_ARG, "null argument");
  OFD_CHECK(precision == OFD_FP64 || precision == OFD_FP32, OFD_ERR_ARG, "bad precision");
  OFD_CHECK(max_cycles >= 1, OFD_ERR_ARG, "max_cycles >= 1");
  Ctx &c = x->c;
  OFD_HIP(hipSetDevice(c.device));
  int rc = ensure_tem_adjoint(c);
  if (rc) return rc;
  rc = ensure_tem_transpose(c);
  if (rc) return rc;
  rc = h2d(c, c.d_lamTb, g, (size_t)c.nnode * sizeof(double));
  if (rc) return rc;
  {
    TemSystem sys(c);
    TemTransposeView view(c);
    rc = solve_core(c, precision, ortho, c.d_lamTb, c.d_lamT, c.d_lamTd, rtol, inner_tol,
                    kspace, max_cycles, iters, cycles, relres);
  }
  if (rc && rc != OFD_ERR_NOCONV) return rc;
  OFD_HIP(hipMemcpyAsync(c.d_field[OFD_LAMT], c.d_lamT, (size_t)c.nnode * sizeof(double),
                         hipMemcpyDeviceToDevice, c.stream));
  OFD_HIP(hipStreamSynchronize(c.stream));
  return rc;
}

int ofd_vjp_tem(ofd_ctx *x, const double *lam, double *gT, double *grT, double *gu,
                double *gks) {
  OFD_CHECK(x && lam, OFD_ERR_ARG, "null argument");
  Ctx &c = x->c;
  int rc = tem_adjoint_ready(c);
  if (rc) return rc;
  const size_t N = (size_t)c.nnode;
  rc = h2d(c, c.d_lamTb, lam, N * sizeof(double));
  if (rc) return rc;
  OFD_HIP(launch_alpha_stage(c));
  // rows T, rT, u (3), then the scale (hexes) or phi, scale (tets and prisms,
  // dual numbers)
  size_t ks_row = 5;
  if (c.nshl != 8) {
    OFD_HIP(launch_vjp_tem_dual(c, c.d_lamTb, c.d_vjpT));
    ks_row = 6;
  } else {
    OFD_HIP(launch_zero(c, c.d_vjpT, N * 6));
    OFD_HIP(launch_vjp_tem(c, c.d_lamTb, c.d_vjpT));
  }
  if (gT) rc = d2h(c, gT, c.d_vjpT, N * sizeof(double));
  if (!rc && grT) rc = d2h(c, grT, c.d_vjpT + N, N * sizeof(double));
  if (!rc && gu) rc = d2h(c, gu, c.d_vjpT + 2 * N, 3 * N * sizeof(double));
  if (!rc && gks) rc = d2h(c, gks, c.d_vjpT + ks_row * N, N * sizeof(double));
  if (rc) return rc;
  OFD_HIP(hipStreamSynchronize(c.stream));
  return OFD_OK;
}

int ofd_jt_apply_tem(ofd_ctx *x, const double *lam, double *y) {
  OFD_CHECK(x && lam && y, OFD_ERR_ARG, "null argument");
  Ctx &c = x->c;
  int rc = tem_adjoint_ready(c);
  if (rc) return rc;
  const size_t N = (size_t)c.nnode;
  rc = h2d(c, c.d_lamTb, lam, N * sizeof(double));
  if (rc) return rc;
  OFD_HIP(launch_alpha_stage(c));
  OFD_HIP(launch_jt_tem(c, c.d_lamTb, c.d_jtT));
  return d2h(c, y, c.d_jtT, N * sizeof(double));
}

int ofd_solve_tem_adjoint(ofd_ctx *x, int precision, int ortho, double rtol,
                          double inner_tol, int kspace, int max_cycles, const double *g,
                          int *iters, int *cycles, double *relres) {
  OFD_CHECK(x && g, OFD_ERR_ARG, "null argument");
  OFD_CHECK(precision == OFD_FP64 || precision == OFD_FP32, OFD_ERR_ARG, "bad precision");
  OFD_CHECK(max_cycles >= 1, OFD_ERR_ARG, "max_cycles >= 1");
  Ctx &c = x->c;
  int rc = tem_adjoint_ready(c);
  if (rc) return rc;
  rc = ensure_tem_transpose(c);
  if (rc) return rc;
  const size_t N = (size_t)c.nnode;
  rc = h2d(c, c.d_lamTb, g, N * sizeof(double));
  if (rc) return rc;
  OFD_HIP(launch_alpha_stage(c));
  OFD_HIP(launch_zero(c, c.d_lamT, (size_t)c.temT_n4 * 4));
  {
    TemSystem sys(c);
    auto apply = [&](const double *v, double *y) -> hipError_t {
      return launch_jt_tem(c, v, y);
    };
    auto prec = [&](const double *v, double *z, int *it) -> int {
      TemTransposeView view(c);
      int cy = 0;
      double ir = 0.0;
      return solve_core(c, precision, ortho, v, z, c.d_lamTd, inner_tol,
                        precision == OFD_FP32 ? 1e-4 : 1e-2, 30, 100, it, &cy, &ir);
    };
    rc = fgmres(c, (size_t)c.temT_n4 * 4, apply, prec, c.d_lamTb, c.d_lamT, c.d_jtT, rtol,
                kspace, max_cycles, iters, cycles, relres);
  }
  if (rc && rc != OFD_ERR_NOCONV) return rc;
  OFD_HIP(hipMemcpyAsync(c.d_field[OFD_LAMT], c.d_lamT, N * sizeof(double),
                         hipMemcpyDeviceToDevice, c.stream));
  OFD_HIP(hipStreamSynchronize(c.stream));
  return rc;
}

int ofd_time_adjoint(ofd_ctx *x, int what, int reps, double *ms) {
  OFD_CHECK(x && ms && reps > 0, OFD_ERR_ARG, "bad argument");
  OFD_CHECK(what >= OFD_TIME_VJP_NS && what <= OFD_TIME_JT_TEM, OFD_ERR_ARG,
            "unknown kernel");
  Ctx &c = x->c;
  const bool tem = what >= OFD_TIME_VJP_TEM;
  int rc = tem ? tem_adjoint_ready(c) : ensure_adjoint(c);
  if (rc) return rc;
  OFD_CHECK(c.nshl == 8 || c.nshl == 4 || c.nshl == 6, OFD_ERR_UNSUPPORTED,
            "the NS adjoint kernels are implemented for trilinear hexes, linear tets and prisms");
  OFD_HIP(hipSetDevice(c.device));
  const size_t N = (size_t)c.nnode;
  OFD_HIP(launch_alpha_stage(c));
  hipEvent_t e0, e1;
  OFD_HIP(hipEventCreate(&e0));
  OFD_HIP(hipEventCreate(&e1));
  for (int r = -1; r < reps; ++r) {  // r = -1: warm-up launch
    if (r == 0) OFD_HIP(hipEventRecord(e0, c.stream));
    if (what == OFD_TIME_VJP_NS) {
      OFD_HIP(launch_zero(c, c.d_vjp, N * 8));
      OFD_HIP(launch_vjp_ns(c, 1, c.d_lamb));
    } else if (what == OFD_TIME_JT_NS) {
      OFD_HIP(launch_jt_ns(c, 1, c.d_lamb, c.d_jt));
    } else if (what == OFD_TIME_VJP_TEM && c.nshl != 8) {
      OFD_HIP(launch_vjp_tem_dual(c, c.d_lamTb, c.d_vjpT));
    } else if (what == OFD_TIME_VJP_TEM) {
      OFD_HIP(launch_zero(c, c.d_vjpT, N * 6));
      OFD_HIP(launch_vjp_tem(c, c.d_lamTb, c.d_vjpT));
    } else {
      OFD_HIP(launch_jt_tem(c, c.d_lamTb, c.d_jtT));
    }
  }
  OFD_HIP(hipEventRecord(e1, c.stream));
  OFD_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  OFD_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = (double)t / reps;
  return OFD_OK;
}

int ofd_set_residual_form(ofd_ctx *x, int form) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  OFD_CHECK(form == 0 || form == 1, OFD_ERR_ARG, "residual form 0 or 1");
  x->c.res_form = form;
  return OFD_OK;
}

int ofd_j_apply(ofd_ctx *x, int subdomain_flag, const double *vu, const double *vp,
                double *yu, double *yp) {
  OFD_CHECK(x && vu && vp && yu && yp, OFD_ERR_ARG, "null argument");
  OFD_CHECK(subdomain_flag >= 1 && subdomain_flag <= 3, OFD_ERR_ARG, "subdomain_flag 1..3");
  Ctx &c = x->c;
  int rc = ensure_adjoint(c);
  if (rc) return rc;
  OFD_CHECK(c.nshl == 8 || c.nshl == 4 || c.nshl == 6, OFD_ERR_UNSUPPORTED,
            "the NS adjoint kernels are implemented for trilinear hexes, linear tets and prisms");
  OFD_HIP(hipSetDevice(c.device));
  rc = put_node4(c, vu, vp, c.d_lamb);
  if (rc) return rc;
  OFD_HIP(launch_alpha_stage(c));
  OFD_HIP(launch_jv_ns(c, subdomain_flag, c.d_lamb, c.d_jt));
  rc = get_node4(c, c.d_jt, yu, yp);
  OFD_HIP(hipStreamSynchronize(c.stream));
  return rc;
}

// ---- exact NS + VOF Jacobian and the coupled quench adjoint ----

static int nsvof_adjoint_ready(Ctx &c) {
  int rc = ensure_nsvof(c);
  if (rc) return rc;
  OFD_CHECK(c.quench, OFD_ERR_ARG, "no quench parameters (ofd_set_quench_params)");
  OFD_CHECK(c.nshl == 8 || c.nshl == 4 || c.nshl == 6, OFD_ERR_UNSUPPORTED,
            "the dual-number kernels are implemented for trilinear hexes, linear tets and prisms");
  OFD_HIP(hipSetDevice(c.device));
  if (c.vadj_n4 != c.nsvV_n4 || !c.d_vjpV) {
    const size_t np = (size_t)c.nsvV_n4 * 4;
    double **bufs[4] = {&c.d_lamV, &c.d_lamVb, &c.d_lamVd, &c.d_jtV};
    for (double **b : bufs) {
      rc = dalloc(c, b, np);
      if (rc) return rc;
      OFD_HIP(hipMemsetAsync(*b, 0, np * sizeof(double), c.stream));
    }
    rc = dalloc(c, &c.d_vjpV, (size_t)c.nnode * 10);
    if (rc) return rc;
    c.vadj_n4 = c.nsvV_n4;
  }
  if (c.t_nsv_icnt != c.icnt) {
    rc = dalloc(c, &c.t_AV, (size_t)c.icnt * 24);
    if (!rc) rc = dalloc(c, &c.t_PV, (size_t)c.nnode * 17);
    if (!rc) rc = dalloc(c, &c.t_AV32, (size_t)0);
    if (!rc) rc = dalloc(c, &c.t_ApV, (size_t)0);
    if (!rc) rc = dalloc(c, &c.t_ApV32, (size_t)0);
    if (rc) return rc;
    c.t_nsv_icnt = c.icnt;
    c.t_matV_valid = false;
  }
  return OFD_OK;
}

// the NS + VOF solver state of the stored matrix swapped with that of its
// transpose, for the inner solves of the adjoint
struct NsvofTransposeView {
  Ctx &c;
  explicit NsvofTransposeView(Ctx &cc) : c(cc) { swap_all(); }
  ~NsvofTransposeView() { swap_all(); }
  void swap_all() {
    std::swap(c.d_AV, c.t_AV);
    std::swap(c.d_AV32, c.t_AV32);
    std::swap(c.d_PV, c.t_PV);
    std::swap(c.d_ApV, c.t_ApV);
    std::swap(c.d_ApV32, c.t_ApV32);
    std::swap(c.matV_valid, c.t_matV_valid);
    std::swap(c.precV_valid, c.t_precV_valid);
    std::swap(c.pV_valid, c.t_pV_valid);
    std::swap(c.aV32_valid, c.t_aV32_valid);
    std::swap(c.apV32_valid, c.t_apV32_valid);
  }
};

int ofd_assemble_nsvof_jac(ofd_ctx *x) {
  OFD_CHECK(x, OFD_ERR_ARG, "null context");
  Ctx &c = x->c;
  int rc = nsvof_adjoint_ready(c);
  if (rc) return rc;
  OFD_HIP(launch_alpha_stage(c, true));
  for (int col = 0; col + 1 < (int)c.color_off.size(); ++col)
    OFD_HIP(launch_nsvof_jac(c, col));
  c.matV_valid = true;
  c.precV_valid = c.pV_valid = c.aV32_valid = c.apV32_valid = false;
  c.t_matV_valid = false;
  OFD_HIP(hipStreamSynchronize(c.stream));
  return OFD_OK;
}

int ofd_vjp_nsvof(ofd_ctx *x, const double *lu, const double *lp, const double *lphi,
                  double *g) {
  OFD_CHECK(x && lu && lp && lphi && g, OFD_ERR_ARG, "null argument");
  Ctx &c = x->c;
  int rc = nsvof_adjoint_ready(c);
  if (rc) return rc;
  const int N = c.nnode;
  std::vector<double> v((size_t)c.nsvV_n4 * 4, 0.0);
  nsvof_pack(N, lu, lp, lphi, v.data());
  rc = h2d(c, c.d_lamVb, v.data(), v.size() * sizeof(double));
  if (rc) return rc;
  OFD_HIP(launch_alpha_stage(c, true));
  OFD_HIP(launch_zero(c, c.d_vjpV, (size_t)N * 10));
  for (int col = 0; col + 1 < (int)c.color_off.size(); ++col)
    OFD_HIP(launch_nsvof_vjp(c, col, c.d_lamVb, c.d_lamVb + (size_t)4 * N, c.d_vjpV));
  rc = d2h(c, g, c.d_vjpV, (size_t)N * 10 * sizeof(double));
  if (rc) return rc;
  OFD_HIP(hipStreamSynchronize(c.stream));
  return OFD_OK;
}

int ofd_vjp_tem_dual(ofd_ctx *x, const double *lam, double *g) {
  OFD_CHECK(x && lam && g, OFD_ERR_ARG, "null argument");
  Ctx &c = x->c;
  int rc = nsvof_adjoint_ready(c);
  if (rc) return rc;
  const size_t N = (size_t)c.nnode;
  rc = h2d(c, c.d_lamVb, lam, N * sizeof(double));
  if (rc) return rc;
  OFD_HIP(launch_alpha_stage(c));
  OFD_HIP(launch_zero(c, c.d_vjpV, N * 7));
  const double *ks = c.ks_on ? c.d_field[OFD_KSCALE] : nullptr;
  for (int col = 0; col + 1 < (int)c.color_off.size(); ++col)
    OFD_HIP(launch_tem_dual_vjp(c, col, ks, c.d_lamVb, c.d_vjpV, nullptr));
  rc = d2h(c, g, c.d_vjpV, N * 7 * sizeof(double));
  if (rc) return rc;
  OFD_HIP(hipStreamSynchronize(c.stream));
  return OFD_OK;
}

int ofd_spmv_nsvof_t(ofd_ctx *x, const double *xu, const double *xp, const double *xl,
                     double *yu, double *yp, double *yl) {
  OFD_CHECK(x && xu && xp && xl && yu && yp && yl, OFD_ERR_ARG, "null argument");
  Ctx &c = x->c;
  int rc = nsvof_adjoint_ready(c);
  if (rc) return rc;
  OFD_CHECK(c.matV_valid, OFD_ERR_ARG, "no assembled NSVOF matrix");
  const int N = c.nnode;
  std::vector<double> v((size_t)c.nsvV_n4 * 4, 0.0);
  nsvof_pack(N, xu, xp, xl, v.data());
  rc = h2d(c, c.d_lamVb, v.data(), v.size() * sizeof(double));
  if (rc) return rc;
  OFD_HIP(launch_spmv_v_t(c, c.d_lamVb, c.d_jtV));
  rc = d2h(c, v.data(), c.d_jtV, v.size() * sizeof(double));
  if (rc) return rc;
  nsvof_unpack(N, v.data(), yu, yp, yl);
  return OFD_OK;
}

int ofd_solve_nsvof_t(ofd_ctx *x, int precision, int ortho, double rtol, double inner_tol,
                      int kspace, int max_cycles, const double *gu, const double *gp,
                      const double *gl, double *lu, double *lp, double *ll, int *iters,
                      int *cycles, double *relres) {
  OFD_CHECK(x && gu && gp && gl && lu && lp && ll, OFD_ERR_ARG, "null argument");
  OFD_CHECK(precision == OFD_FP64 || precision == OFD_FP32, OFD_ERR_ARG, "bad precision");
  OFD_CHECK(max_cycles >= 1, OFD_ERR_ARG, "max_cycles >= 1");
  Ctx &c = x->c;
  int rc = nsvof_adjoint_ready(c);
  if (rc) return rc;
  OFD_CHECK(c.matV_valid, OFD_ERR_ARG, "no assembled NSVOF matrix");
  const int N = c.nnode;
  if (!c.t_matV_valid) {
    OFD_HIP(launch_nsvof_transpose(c, c.t_AV));
    c.t_matV_valid = true;
    c.t_precV_valid = c.t_pV_valid = c.t_aV32_valid = c.t_apV32_valid = false;
  }
  std::vector<double> v((size_t)c.nsvV_n4 * 4, 0.0);
  nsvof_pack(N, gu, gp, gl, v.data());
  rc = h2d(c, c.d_lamVb, v.data(), v.size() * sizeof(double));
  if (rc) return rc;
  OFD_HIP(launch_zero(c, c.d_lamV, v.size()));
  NsvofSystem sys(c);
  auto apply = [&](const double *vin, double *y) -> hipError_t {
    return launch_spmv_v_t(c, vin, y);
  };
  auto prec = [&](const double *vin, double *z, int *it) -> int {
    int cy = 0;
    double ir = 0.0;
    const double ctol = precision == OFD_FP32 ? 1e-4 : 1e-2;
    NsvofTransposeView view(c);
    return solve_core(c, precision, ortho, vin, z, c.d_lamVd, inner_tol, ctol, 30, 100, it,
                      &cy, &ir);
  };
  rc = fgmres(c, v.size(), apply, prec, c.d_lamVb, c.d_lamV, c.d_jtV, rtol, kspace, max_cycles,
              iters, cycles, relres);
  if (rc && rc != OFD_ERR_NOCONV) return rc;
  const int rc2 = d2h(c, v.data(), c.d_lamV, v.size() * sizeof(double));
  if (rc2) return rc2;
  nsvof_unpack(N, v.data(), lu, lp, ll);
  return rc;
}

int ofd_ctx_bytes(ofd_ctx *x, int64_t *bytes) {
  OFD_CHECK(x && bytes, OFD_ERR_ARG, "null argument");
  *bytes = x->c.dev_bytes;
  return OFD_OK;
}

}  // extern "C"
