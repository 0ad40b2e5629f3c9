// genSparStruc (genSparStruc.f90:4-105) on the device, with the maps the
// kernels need beside it.
//
//   node -> element incidence   one thread per element adds itself to the
//                               count of each of its nodes; an inclusive scan
//                               gives the offsets; a second pass fills the
//                               lists (their order does not matter: they are
//                               merged and sorted below)
//   row of node i               one thread per node merges the nodes of its
//                               elements into a sorted list without repeats,
//                               at most 27 NSHL long (the reference's Asadj
//                               bound); a first pass counts, a scan gives
//                               col, a second pass writes row and the
//                               diagonal position (SparseGMRES_NS.f90:53-61)
//   tmap                        per nonzero (i, j), the position of (j, i)
//                               (the pattern is symmetric), one thread per
//                               nonzero
//   emap                        per element and node pair (a, b), the nonzero
//                               of (ien[a], ien[b]): SparseMatLoc_3D's search
//                               done once, at e*64 + a + 8 b
//
// Every output is a function of the mesh alone, whatever the order of the
// integer atomics, and equals the reference's col/row (sorted columns per
// row).  Errors (a row past 27 NSHL, an asymmetric pattern) set a flag.
#include <hipcub/hipcub.hpp>

#include "ofd_internal.h"

namespace ofd {

namespace {

constexpr int kAdjCap = 27 * 8;  // 27 NSHL for the largest element, the hex

__global__ void k_node_elem_count(int E, int S, const int32_t *__restrict__ ien,
                                  int32_t *__restrict__ cnt) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  for (int a = 0; a < S; ++a) atomicAdd(&cnt[ien[(size_t)a * E + e]], 1);
}

__global__ void k_node_elem_fill(int E, int S, const int32_t *__restrict__ ien,
                                 const int32_t *__restrict__ off, int32_t *__restrict__ pos,
                                 int32_t *__restrict__ inc) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  for (int a = 0; a < S; ++a) {
    const int n = ien[(size_t)a * E + e];
    const int p = atomicAdd(&pos[n], 1);
    inc[off[n] + p] = e;
  }
}

// sorted, repeat-free neighbour list of node i; col == nullptr: count only
__global__ void __launch_bounds__(256)
k_adjacency(int N, int E, int S, const int32_t *__restrict__ ien,
            const int32_t *__restrict__ off, const int32_t *__restrict__ inc,
            const int32_t *__restrict__ col, int32_t *__restrict__ out,
            int32_t *__restrict__ diag, int32_t *__restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int cap = 27 * S;
  int32_t L[kAdjCap];
  int len = 0;
  for (int p = off[i]; p < off[i + 1]; ++p) {
    const int e = inc[p];
    for (int k = 0; k < S; ++k) {
      const int kn = ien[(size_t)k * E + e];
      int lo = 0, hi = len;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (L[mid] < kn) lo = mid + 1;
        else hi = mid;
      }
      if (lo < len && L[lo] == kn) continue;
      if (len >= cap) {
        atomicOr(err, 1);
        return;
      }
      for (int t = len; t > lo; --t) L[t] = L[t - 1];
      L[lo] = kn;
      ++len;
    }
  }
  if (!col) {
    out[i] = len;
    return;
  }
  int32_t *row = out + col[i];
  for (int t = 0; t < len; ++t) {
    row[t] = L[t];
    if (L[t] == i) diag[i] = col[i] + t;
  }
}

__device__ __forceinline__ int lower_bound(const int32_t *__restrict__ row, int lo, int hi,
                                           int goal) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (row[mid] < goal) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// one thread per nonzero k: its row i is the last with col[i] <= k
__global__ void k_tmap(int N, size_t icnt, const int32_t *__restrict__ col,
                       const int32_t *__restrict__ row, int32_t *__restrict__ tmap,
                       int32_t *__restrict__ err) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= icnt) return;
  int lo = 0, hi = N;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if ((size_t)col[mid] <= k) lo = mid + 1;
    else hi = mid;
  }
  const int i = lo - 1;
  const int j = row[k];
  const int p = lower_bound(row, col[j], col[j + 1], i);
  if (p == col[j + 1] || row[p] != i) {
    atomicOr(err, 2);
    tmap[k] = 0;
  } else {
    tmap[k] = p;
  }
}

__global__ void k_emap(int E, int S, const int32_t *__restrict__ ien,
                       const int32_t *__restrict__ col, const int32_t *__restrict__ row,
                       int32_t *__restrict__ emap) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)E * 64) return;
  const int e = (int)(t >> 6), a = (int)(t & 7), b = (int)((t >> 3) & 7);
  if (a >= S || b >= S) {
    emap[t] = 0;
    return;
  }
  const int ia = ien[(size_t)a * E + e], ib = ien[(size_t)b * E + e];
  emap[t] = lower_bound(row, col[ia], col[ia + 1], ib);
}

// First touch of each nonzero in the assembly's launch order (colour by
// colour, elements of a colour share no node, hence no nonzero): bit
// a + 8 b of first[e] is set when element e is the first to add into the
// nonzero of its pair (a, b).  One wavefront per element of one colour;
// launched colour by colour, so the result does not depend on timing.
__global__ void __launch_bounds__(256)
k_first_touch(int S, const int32_t *__restrict__ elems, int first_idx, int count,
              const int32_t *__restrict__ emap, uint8_t *__restrict__ touched,
              uint64_t *__restrict__ first) {
  const int lane = threadIdx.x & 63;
  const int idx = first_idx + blockIdx.x * 4 + (threadIdx.x >> 6);
  if (idx >= first_idx + count) return;  // uniform over the wave
  const int e = elems[idx];
  const int a = lane & 7, b = lane >> 3;
  bool mine = false;
  if (a < S && b < S) {
    const int k = emap[(size_t)e * 64 + lane];
    mine = touched[k] == 0;
    touched[k] = 1;
  }
  const unsigned long long m = __ballot(mine);
  if (lane == 0) first[e] = m;
}

// The scatter map and first-touch masks in colour order (position t of
// the colour-sorted element list), and each element's position: the hex NS
// pair kernel reads them by position, with no element index in between.
__global__ void __launch_bounds__(256)
k_colour_maps(int E, const int32_t *__restrict__ elems, const int32_t *__restrict__ emap,
              const uint64_t *__restrict__ first, int32_t *__restrict__ emap_c,
              uint64_t *__restrict__ first_c, int32_t *__restrict__ cpos) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)E * 64) return;
  const int t = (int)(i >> 6), j = (int)(i & 63);
  const int e = elems[t];
  emap_c[i] = emap[(size_t)e * 64 + j];
  if (j == 0) {
    first_c[t] = first[e];
    cpos[e] = t;
  }
}

}  // namespace

hipError_t launch_colour_maps(Ctx &c) {
  const size_t n = (size_t)c.nelem * 64;
  hipLaunchKernelGGL(k_colour_maps, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream,
                     c.nelem, c.d_color_elems, c.d_emap, c.d_first, c.d_emap_c, c.d_first_c,
                     c.d_cpos);
  return hipGetLastError();
}

hipError_t launch_first_touch(Ctx &c, uint8_t *touched) {
  for (int col = 0; col + 1 < (int)c.color_off.size(); ++col) {
    const int f = c.color_off[col], n = c.color_off[col + 1] - f;
    if (n <= 0) continue;
    hipLaunchKernelGGL(k_first_touch, dim3((n + 3) / 4), dim3(256), 0, c.stream, c.nshl,
                       c.d_color_elems, f, n, c.d_emap, touched, c.d_first);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_node_elem_count(Ctx &c, int32_t *cnt) {
  hipLaunchKernelGGL(k_node_elem_count, dim3((c.nelem + 255) / 256), dim3(256), 0, c.stream,
                     c.nelem, c.nshl, c.d_ien, cnt);
  return hipGetLastError();
}

hipError_t launch_node_elem_fill(Ctx &c, const int32_t *off, int32_t *pos, int32_t *inc) {
  hipLaunchKernelGGL(k_node_elem_fill, dim3((c.nelem + 255) / 256), dim3(256), 0, c.stream,
                     c.nelem, c.nshl, c.d_ien, off, pos, inc);
  return hipGetLastError();
}

hipError_t launch_adjacency(Ctx &c, const int32_t *off, const int32_t *inc, const int32_t *col,
                            int32_t *out, int32_t *diag, int32_t *err) {
  hipLaunchKernelGGL(k_adjacency, dim3((c.nnode + 255) / 256), dim3(256), 0, c.stream, c.nnode,
                     c.nelem, c.nshl, c.d_ien, off, inc, col, out, diag, err);
  return hipGetLastError();
}

hipError_t launch_tmap_emap(Ctx &c, int32_t *err) {
  const size_t nz = (size_t)c.icnt;
  hipLaunchKernelGGL(k_tmap, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, c.stream,
                     c.nnode, nz, c.d_col, c.d_row, c.d_tmap, err);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const size_t nt = (size_t)c.nelem * 64;
  hipLaunchKernelGGL(k_emap, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c.stream,
                     c.nelem, c.nshl, c.d_ien, c.d_col, c.d_row, c.d_emap);
  return hipGetLastError();
}

// out[0] = 0, out[1..n] = inclusive prefix sums of in[0..n); tmp/tmp_bytes
// as hipcub takes them (tmp == nullptr: the size only)
hipError_t scan_offsets(Ctx &c, const int32_t *in, int32_t *out, int n, void *tmp,
                        size_t &tmp_bytes) {
  if (tmp) {
    hipError_t e = hipMemsetAsync(out, 0, sizeof(int32_t), c.stream);
    if (e != hipSuccess) return e;
  }
  return hipcub::DeviceScan::InclusiveSum(tmp, tmp_bytes, in, out + 1, n, c.stream);
}

}  // namespace ofd
