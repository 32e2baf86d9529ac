// Face connectivity on the device: the interface and boundary lists of
// P4estMesh{3} (Trixi's P4estInterfaceContainer / P4estBoundaryContainer,
// which p4est's connectivity fills on the host) and of
// P4estMeshCubedSphere2D (src/meshes/p4est_cubed_sphere.jl:82-299), built
// from the node coordinates alone.  The same rules as the host builders of
// mesh3d.cpp / mesh2d.cpp, so the lists are equal entry for entry:
//  * every element face gets a key, its centre quantised on a grid of 1e-5
//    of the mesh diameter (3D; wrapped into the period on periodic axes) or
//    1e-6 of the sphere radius (2D);
//  * a radix sort of the keys (stable, so equal keys keep ascending
//    element-face ids) puts the two sides of an interface next to each
//    other; the side with the smaller id is the primary, indexed
//    (:i_forward, :j_forward); a face alone is a boundary (3D) or an error
//    (2D, a closed sphere);
//  * interfaces and boundaries are numbered in ascending primary id by an
//    exclusive scan, the host's sort order;
//  * the secondary side's node-index codes are the orientation under which
//    every surface node coincides within 1e-8 of the diameter (3D, eight
//    candidates) or 1e-9 of the radius (2D, forward or backward).
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "atmo_internal.hpp"
#include "handle.hpp"

namespace atmo {

namespace {

// mesh extent and the derived grid spacing and tolerance
struct Extent {
  double lo[3], hi[3];
  double quant, tol;
};

// node-index codes (the host's kBegin ... kJBwd)
__device__ __forceinline__ int code_index(int code, int a, int b, int n) {
  switch (code) {
    case kBegin: return 0;
    case kEnd: return n - 1;
    case kIFwd: return a;
    case kIBwd: return n - 1 - a;
    case kJFwd: return b;
    default: return n - 1 - b;
  }
}

__device__ __forceinline__ void face_codes(int face, int c1, int c2, int out[3]) {
  const int d = face / 2;
  int r = 0;
  for (int q = 0; q < 3; ++q) {
    if (q == d) out[q] = (face % 2 == 0) ? kBegin : kEnd;
    else out[q] = (r++ == 0) ? c1 : c2;
  }
}

__device__ __forceinline__ const double* node3(const double* X, int n, int i, int j, int k,
                                               int e) {
  return X + 3 * ((size_t)i + n * ((size_t)j + n * ((size_t)k + n * (size_t)e)));
}

// key of a quantised centre, counted from the lower corner of the mesh: 21
// bits per component (grid indices up to 2e6, the 2D sphere's diameter over
// its spacing), shifted by 16 to keep a rounded -0 or -1 in range
__device__ __forceinline__ unsigned long long pack_key(const long long k[3]) {
  unsigned long long key = 0;
  for (int c = 0; c < 3; ++c)
    key = (key << 21) | ((unsigned long long)(k[c] + 16) & ((1ULL << 21) - 1));
  return key;
}

// per-block minimum and maximum of the coordinates, [blocks][6]
__global__ __launch_bounds__(256) void extent_partial(const double* __restrict__ X, long long nodes,
                                                      double* __restrict__ partial) {
  __shared__ double s[6][256];
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < nodes;
       q += (long long)gridDim.x * 256)
    for (int c = 0; c < 3; ++c) {
      lo[c] = fmin(lo[c], X[3 * q + c]);
      hi[c] = fmax(hi[c], X[3 * q + c]);
    }
  for (int c = 0; c < 3; ++c) {
    s[c][threadIdx.x] = lo[c];
    s[3 + c][threadIdx.x] = hi[c];
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int c = 0; c < 3; ++c) {
        s[c][threadIdx.x] = fmin(s[c][threadIdx.x], s[c][threadIdx.x + w]);
        s[3 + c][threadIdx.x] = fmax(s[3 + c][threadIdx.x], s[3 + c][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) partial[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

// the extent from the partials, then the diameter's grid and tolerance (3D)
// or the radius's (2D: the first node lies on the sphere)
__global__ void extent_final(const double* __restrict__ partial, int blocks,
                             const double* __restrict__ X, int dim, Extent* ext) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0) return;
  Extent x;
  for (int c = 0; c < 3; ++c) {
    x.lo[c] = 1e300;
    x.hi[c] = -1e300;
  }
  for (int b = 0; b < blocks; ++b)
    for (int c = 0; c < 3; ++c) {
      x.lo[c] = fmin(x.lo[c], partial[b * 6 + c]);
      x.hi[c] = fmax(x.hi[c], partial[b * 6 + 3 + c]);
    }
  if (dim == 3) {
    const double diam = sqrt((x.hi[0] - x.lo[0]) * (x.hi[0] - x.lo[0]) +
                             (x.hi[1] - x.lo[1]) * (x.hi[1] - x.lo[1]) +
                             (x.hi[2] - x.lo[2]) * (x.hi[2] - x.lo[2]));
    x.quant = 1e-5 * diam;
    x.tol = 1e-8 * diam;
  } else {
    const double R = sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
    x.quant = 1e-6 * R;
    x.tol = 1e-9 * R;
  }
  *ext = x;
}

struct Periods {
  double p[3];  // 0: not periodic
};

// keys of the 6 faces of every hexahedron: the mean of the four corner nodes
// (in the host's order), wrapped into [0, period) on periodic axes; the grid
// counts from the lower corner of the mesh (from 0 on periodic axes)
__global__ void face_keys_3d(const double* __restrict__ X, int n, int E, Periods per,
                             const Extent* __restrict__ ext, unsigned long long* __restrict__ key,
                             int* __restrict__ id) {
#pragma clang fp contract(off)
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 6 * E) return;
  const int e = q / 6, f = q % 6;
  int codes[3];
  face_codes(f, kIFwd, kJFwd, codes);
  const int cs[4][2] = {{0, 0}, {n - 1, 0}, {0, n - 1}, {n - 1, n - 1}};
  double ctr[3] = {0.0, 0.0, 0.0};
  for (int t = 0; t < 4; ++t) {
    const int a = cs[t][0], b = cs[t][1];
    const double* p = node3(X, n, code_index(codes[0], a, b, n), code_index(codes[1], a, b, n),
                            code_index(codes[2], a, b, n), e);
    for (int c = 0; c < 3; ++c) ctr[c] += 0.25 * p[c];
  }
  long long k[3];
  for (int c = 0; c < 3; ++c) {
    double v = ctr[c];
    double origin = ext->lo[c];
    if (per.p[c] > 0) {
      v -= per.p[c] * floor(v / per.p[c] + 1e-9);
      origin = 0.0;
    }
    k[c] = llround((v - origin) / ext->quant);
  }
  key[q] = pack_key(k);
  id[q] = q;
}

// keys of the 4 edges of every quadrilateral of the cubed sphere: the mean of
// the edge's end nodes
__global__ void face_keys_2d(const double* __restrict__ X, int n, int E,
                             const Extent* __restrict__ ext, unsigned long long* __restrict__ key,
                             int* __restrict__ id) {
#pragma clang fp contract(off)
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 4 * E) return;
  const int e = q / 4, f = q % 4;
  auto node = [&](int k) {
    const int i = f == 0 ? 0 : (f == 1 ? n - 1 : k);
    const int j = f == 2 ? 0 : (f == 3 ? n - 1 : k);
    return X + 3 * ((size_t)i + n * ((size_t)j + n * (size_t)e));
  };
  const double* p0 = node(0);
  const double* p1 = node(n - 1);
  long long k[3];
  for (int c = 0; c < 3; ++c) k[c] = llround((0.5 * (p0[c] + p1[c]) - ext->lo[c]) / ext->quant);
  key[q] = pack_key(k);
  id[q] = q;
}

// roles from the sorted keys: 1 primary (partner = the next id), 2 secondary,
// 0 alone; status 1 where more than two faces share a key
__global__ void face_roles(const unsigned long long* __restrict__ key,
                           const int* __restrict__ id, int nf, int* __restrict__ primary,
                           int* __restrict__ alone, int* __restrict__ partner,
                           int* __restrict__ status) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nf) return;
  const unsigned long long k = key[s];
  const bool prev = s > 0 && key[s - 1] == k;
  const bool next = s + 1 < nf && key[s + 1] == k;
  const int me = id[s];
  if (prev && next) atomicMax(status, 1);
  primary[me] = next && !prev;
  alone[me] = !prev && !next;
  partner[me] = next ? id[s + 1] : (prev ? id[s - 1] : -1);
}

__device__ __forceinline__ double dist3(const double* a, const double* b, const Periods& per) {
#pragma clang fp contract(off)
  double s = 0.0;
  for (int c = 0; c < 3; ++c) {
    double dd = a[c] - b[c];
    if (per.p[c] > 0) dd -= per.p[c] * rint(dd / per.p[c]);
    s += dd * dd;
  }
  return sqrt(s);
}

// one interface per primary face: element ids and node-index codes
__global__ void interfaces_3d(const double* __restrict__ X, int n, int E, Periods per,
                              const Extent* __restrict__ ext, const int* __restrict__ primary,
                              const int* __restrict__ partner, const int* __restrict__ slot,
                              int32_t* __restrict__ ids, int32_t* __restrict__ codes,
                              int* __restrict__ status) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 6 * E || !primary[q]) return;
  const int ep = q / 6, fp = q % 6, o = partner[q], es = o / 6, fs = o % 6;
  const int opts[8][2] = {{kIFwd, kJFwd}, {kIBwd, kJFwd}, {kIFwd, kJBwd}, {kIBwd, kJBwd},
                          {kJFwd, kIFwd}, {kJBwd, kIFwd}, {kJFwd, kIBwd}, {kJBwd, kIBwd}};
  int cp[3], cs[3];
  face_codes(fp, kIFwd, kJFwd, cp);
  const double tol = ext->tol;
  bool found = false;
  for (int t = 0; t < 8 && !found; ++t) {
    face_codes(fs, opts[t][0], opts[t][1], cs);
    bool ok = true;
    for (int b = 0; b < n && ok; ++b)
      for (int a = 0; a < n && ok; ++a) {
        const double* pp = node3(X, n, code_index(cp[0], a, b, n), code_index(cp[1], a, b, n),
                                 code_index(cp[2], a, b, n), ep);
        const double* ps = node3(X, n, code_index(cs[0], a, b, n), code_index(cs[1], a, b, n),
                                 code_index(cs[2], a, b, n), es);
        if (dist3(pp, ps, per) > tol) ok = false;
      }
    found = ok;
  }
  if (!found) {
    atomicMax(status, 2);
    return;
  }
  const int i = slot[q];
  ids[2 * i] = ep;
  ids[2 * i + 1] = es;
  for (int c = 0; c < 3; ++c) {
    codes[6 * i + c] = cp[c];
    codes[6 * i + 3 + c] = cs[c];
  }
}

// one boundary per lone face: element, node-index codes and kind (the -z face
// of a shell column is :inside (0), +z :outside (1), others 2)
__global__ void boundaries_3d(int E, const int* __restrict__ alone, const int* __restrict__ slot,
                              int32_t* __restrict__ element, int32_t* __restrict__ codes,
                              int32_t* __restrict__ kind) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 6 * E || !alone[q]) return;
  const int e = q / 6, f = q % 6, i = slot[q];
  int cb[3];
  face_codes(f, kIFwd, kJFwd, cb);
  element[i] = e;
  for (int c = 0; c < 3; ++c) codes[3 * i + c] = cb[c];
  kind[i] = f == 4 ? 0 : (f == 5 ? 1 : 2);
}

__device__ __forceinline__ void edge_codes(int f, int surf, int32_t out[2]) {
  if (f == 0) { out[0] = kBegin; out[1] = surf; }
  else if (f == 1) { out[0] = kEnd; out[1] = surf; }
  else if (f == 2) { out[0] = surf; out[1] = kBegin; }
  else { out[0] = surf; out[1] = kEnd; }
}

__global__ void interfaces_2d(const double* __restrict__ X, int n, int E,
                              const Extent* __restrict__ ext, const int* __restrict__ primary,
                              const int* __restrict__ alone, const int* __restrict__ partner,
                              const int* __restrict__ slot, int32_t* __restrict__ ids,
                              int32_t* __restrict__ codes, int* __restrict__ status) {
#pragma clang fp contract(off)
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 4 * E) return;
  if (alone[q]) {
    atomicMax(status, 3);
    return;
  }
  if (!primary[q]) return;
  const int ep = q / 4, fp = q % 4, o = partner[q], es = o / 4, fs = o % 4;
  auto node = [&](int e, int f, int k) {
    const int i = f == 0 ? 0 : (f == 1 ? n - 1 : k);
    const int j = f == 2 ? 0 : (f == 3 ? n - 1 : k);
    return X + 3 * ((size_t)i + n * ((size_t)j + n * (size_t)e));
  };
  auto d = [](const double* a, const double* c) {
    return sqrt((a[0] - c[0]) * (a[0] - c[0]) + (a[1] - c[1]) * (a[1] - c[1]) +
                (a[2] - c[2]) * (a[2] - c[2]));
  };
  const double tol = ext->tol;
  bool fwd = true;
  for (int k = 0; k < n && fwd; ++k) fwd = !(d(node(ep, fp, k), node(es, fs, k)) > tol);
  if (!fwd)
    for (int k = 0; k < n; ++k)
      if (d(node(ep, fp, k), node(es, fs, n - 1 - k)) > tol) {
        atomicMax(status, 2);
        return;
      }
  const int i = slot[q];
  int32_t cp[2], cs[2];
  edge_codes(fp, kIFwd, cp);
  edge_codes(fs, fwd ? kIFwd : kIBwd, cs);
  ids[2 * i] = ep;
  ids[2 * i + 1] = es;
  codes[4 * i] = cp[0];
  codes[4 * i + 1] = cp[1];
  codes[4 * i + 2] = cs[0];
  codes[4 * i + 3] = cs[1];
}

// counts = exclusive sum at the end + the last flag
__global__ void flag_count(const int* __restrict__ flags, const int* __restrict__ scan, int nf,
                           int* __restrict__ count) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *count = scan[nf - 1] + flags[nf - 1];
}

// device buffers of one connectivity build, released on every return path
struct Scratch {
  std::vector<void*> ptrs;
  ~Scratch() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  bool get(T** p, size_t count) {
    void* q = nullptr;
    if (hipMalloc(&q, count * sizeof(T) + 16) != hipSuccess) return false;
    ptrs.push_back(q);
    *p = static_cast<T*>(q);
    return true;
  }
};

// shared part: keys, sort, roles, scans; leaves primary/alone flags, partner
// and slot (the interface or boundary index of each face) on the device
int build_faces(int dim, int n, int E, const double* X, const Periods& per, Scratch& sc,
                Extent** ext, int** primary, int** alone, int** partner, int** islot,
                int** bslot, int** counts, int** status, hipStream_t s) {
  const int F = dim == 3 ? 6 : 4;
  const int nf = F * E;
  const long long nodes = (long long)E * n * n * (dim == 3 ? n : 1);
  const int blocks = (int)std::min<long long>(1024, (nodes + 255) / 256);
  double* partial = nullptr;
  unsigned long long *key = nullptr, *key_sorted = nullptr;
  int *id = nullptr, *id_sorted = nullptr;
  if (!sc.get(&partial, 6 * (size_t)blocks) || !sc.get(ext, 1) || !sc.get(&key, nf) ||
      !sc.get(&key_sorted, nf) || !sc.get(&id, nf) || !sc.get(&id_sorted, nf) ||
      !sc.get(primary, nf) || !sc.get(alone, nf) || !sc.get(partner, nf) || !sc.get(islot, nf) ||
      !sc.get(bslot, nf) || !sc.get(counts, 2) || !sc.get(status, 1))
    return fail(nullptr, ATMO_ERR_DEVICE, "device allocation failed");
  if (hipMemsetAsync(*status, 0, sizeof(int), s) != hipSuccess)
    return fail(nullptr, ATMO_ERR_DEVICE, "device memset failed");
  hipLaunchKernelGGL(extent_partial, dim3(blocks), dim3(256), 0, s, X, nodes, partial);
  hipLaunchKernelGGL(extent_final, dim3(1), dim3(64), 0, s, partial, blocks, X, dim, *ext);
  const dim3 g((nf + 255) / 256), b(256);
  if (dim == 3)
    hipLaunchKernelGGL(face_keys_3d, g, b, 0, s, X, n, E, per, *ext, key, id);
  else
    hipLaunchKernelGGL(face_keys_2d, g, b, 0, s, X, n, E, *ext, key, id);
  size_t bytes = 0, bytes_scan = 0;
  if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, key, key_sorted, id, id_sorted, nf, 0,
                                         63, s) != hipSuccess ||
      hipcub::DeviceScan::ExclusiveSum(nullptr, bytes_scan, *primary, *islot, nf, s) !=
          hipSuccess)
    return fail(nullptr, ATMO_ERR_DEVICE, "radix sort / scan set-up failed");
  void* tmp = nullptr;
  if (!sc.get(reinterpret_cast<char**>(&tmp), std::max(bytes, bytes_scan)))
    return fail(nullptr, ATMO_ERR_DEVICE, "device allocation failed");
  const size_t tmp_bytes = std::max(bytes, bytes_scan);
  bytes = tmp_bytes;
  if (hipcub::DeviceRadixSort::SortPairs(tmp, bytes, key, key_sorted, id, id_sorted, nf, 0, 63,
                                         s) != hipSuccess)
    return fail(nullptr, ATMO_ERR_DEVICE, "radix sort failed");
  hipLaunchKernelGGL(face_roles, g, b, 0, s, key_sorted, id_sorted, nf, *primary, *alone,
                     *partner, *status);
  bytes = tmp_bytes;
  if (hipcub::DeviceScan::ExclusiveSum(tmp, bytes, *primary, *islot, nf, s) != hipSuccess)
    return fail(nullptr, ATMO_ERR_DEVICE, "scan failed");
  bytes = tmp_bytes;
  if (hipcub::DeviceScan::ExclusiveSum(tmp, bytes, *alone, *bslot, nf, s) != hipSuccess)
    return fail(nullptr, ATMO_ERR_DEVICE, "scan failed");
  hipLaunchKernelGGL(flag_count, dim3(1), dim3(64), 0, s, *primary, *islot, nf, *counts);
  hipLaunchKernelGGL(flag_count, dim3(1), dim3(64), 0, s, *alone, *bslot, nf, *counts + 1);
  return ATMO_OK;
}

int finish(int* d_status, int* d_counts, int32_t* n_if, int32_t* n_bd, hipStream_t s) {
  int st = 0, cnt[2] = {0, 0};
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(&st, d_status, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(cnt, d_counts, 2 * sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(nullptr, ATMO_ERR_DEVICE, std::string("HIP: ") + hipGetErrorString(e));
  if (st == 1) return fail(nullptr, ATMO_ERR_MESH, "face matched more than two elements");
  if (st == 2) return fail(nullptr, ATMO_ERR_MESH, "no orientation matches the face nodes");
  if (st == 3) return fail(nullptr, ATMO_ERR_MESH, "cubed-sphere edge without exactly two elements");
  if (n_if) *n_if = cnt[0];
  if (n_bd) *n_bd = cnt[1];
  return ATMO_OK;
}

}  // namespace

}  // namespace atmo

using namespace atmo;

extern "C" {

int atmo_mesh3d_connectivity_device(int polydeg, int32_t n_elements,
                                    const double* node_coordinates, const double* period,
                                    int32_t* interface_neighbor_ids,
                                    int32_t* interface_node_indices, int32_t* boundary_element,
                                    int32_t* boundary_node_indices, int32_t* boundary_kind,
                                    int32_t* n_interfaces, int32_t* n_boundaries) {
  if (polydeg < 1 || polydeg > 15 || n_elements < 1 || !node_coordinates ||
      !interface_neighbor_ids || !interface_node_indices || !boundary_element ||
      !boundary_node_indices || !boundary_kind || !n_interfaces || !n_boundaries)
    return fail(nullptr, ATMO_ERR_INVALID, "bad arguments");
  if (n_elements > (1 << 28)) return fail(nullptr, ATMO_ERR_INVALID, "too many elements");
  Periods per{};
  for (int c = 0; c < 3; ++c) per.p[c] = period ? period[c] : 0.0;
  const int n = polydeg + 1, E = n_elements;
  hipStream_t s = nullptr;
  Scratch sc;
  Extent* ext;
  int *primary, *alone, *partner, *islot, *bslot, *counts, *status;
  int rc = build_faces(3, n, E, node_coordinates, per, sc, &ext, &primary, &alone, &partner,
                       &islot, &bslot, &counts, &status, s);
  if (rc != ATMO_OK) return rc;
  const dim3 g((6 * E + 255) / 256), b(256);
  hipLaunchKernelGGL(interfaces_3d, g, b, 0, s, node_coordinates, n, E, per, ext, primary,
                     partner, islot, interface_neighbor_ids, interface_node_indices, status);
  hipLaunchKernelGGL(boundaries_3d, g, b, 0, s, E, alone, bslot, boundary_element,
                     boundary_node_indices, boundary_kind);
  return finish(status, counts, n_interfaces, n_boundaries, s);
}

int atmo_mesh2d_connectivity_device(int polydeg, int32_t n_elements,
                                    const double* node_coordinates,
                                    int32_t* interface_neighbor_ids,
                                    int32_t* interface_node_indices, int32_t* n_interfaces) {
  if (polydeg < 1 || polydeg > 15 || n_elements < 1 || !node_coordinates ||
      !interface_neighbor_ids || !interface_node_indices || !n_interfaces)
    return fail(nullptr, ATMO_ERR_INVALID, "bad arguments");
  if (n_elements > (1 << 28)) return fail(nullptr, ATMO_ERR_INVALID, "too many elements");
  const int n = polydeg + 1, E = n_elements;
  hipStream_t s = nullptr;
  Scratch sc;
  Extent* ext;
  int *primary, *alone, *partner, *islot, *bslot, *counts, *status;
  int rc = build_faces(2, n, E, node_coordinates, Periods{}, sc, &ext, &primary, &alone,
                       &partner, &islot, &bslot, &counts, &status, s);
  if (rc != ATMO_OK) return rc;
  const dim3 g((4 * E + 255) / 256), b(256);
  hipLaunchKernelGGL(interfaces_2d, g, b, 0, s, node_coordinates, n, E, ext, primary, alone,
                     partner, islot, interface_neighbor_ids, interface_node_indices, status);
  return finish(status, counts, n_interfaces, nullptr, s);
}

}  // extern "C"
