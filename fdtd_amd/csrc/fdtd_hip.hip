// libfdtd_hip: fused Yee half-step kernels for MI355X (gfx950) behind the C
// ABI declared in include/fdtd_hip.h.
//
// One E half step is, per cell and in this order (fdtd/grid.py:275-299):
//   curl = backward-difference curl of H                 (grid.py:54-76)
//   E   += (S * inv_eps) * curl                          (grid.py:283)
//   objects in registration order                        (objects.py:118-269)
//   boundaries in registration order: CPML correction    (boundaries.py:409-459)
//     or periodic plane copy                             (boundaries.py:184-219)
//   sources in registration order, then detectors        (sources.py, detectors.py)
// and the H half step mirrors it (grid.py:301-325). Every floating-point
// operation is the one the reference performs, in the same order, with
// contraction disabled (-ffp-contract=off), so float64 results match the
// numpy reference bit for bit.
//
// Kernels
//   k_main_E / k_main_H  curl, material update, objects and every CPML face
//                        registered before the first periodic boundary; a
//                        workgroup marches along x over a (y,z) tile.
//                        Memory-bound (~0.16 flop/B).
//   k_face_E / k_face_H  one CPML face over its slab, for faces registered
//                        after a periodic boundary (order matters there).
//   k_periodic           plane copy of a periodic boundary.
//   k_tail               one workgroup: sources (in order) then detectors.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fdtd_hip.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define FDTD_HIP(expr)                                                            \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) return fail(-3, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------- device plan
struct FaceDev {
  int32_t axis, high, t, pad;
  const double* bE;
  const double* cE;
  const double* bH;
  const double* cH;
  void* psiE;
  void* psiH;
  int64_t slab;  // cells in the face slab
};

struct ObjDev {
  int32_t kind, x0, x1, y0, y1, z0, z1, pad;
  const void* inv;
  const void* absorb;
};

struct SrcDev {
  int32_t kind, n, compE, compH;
  const int64_t* cells;
  const double* profile;
  double amplitude;
};

struct DetDev {
  int32_t n, pad;
  const int64_t* cells;
  void* bufE;
  void* bufH;
  int64_t slot0E, slot0H;  // slot written at half-step counter c is c - slot0
};

struct Dims {
  int32_t nx, ny, nz, pad;
  int64_t cs;  // component stride nx*ny*nz
};

constexpr int FOLDED_POINTS = 4;

template <typename T>
struct MainArgs {
  Dims d;
  T sc;
  T* F;          // field updated in this half step (E or H)
  const T* G;    // field whose curl drives it (H or E)
  const T* mat;  // inv_eps or inv_mu
  const FaceDev* faces;
  int32_t nfaces;
  const ObjDev* objs;
  int32_t nobjs;
  int32_t bz_shift;  // block extent along z is 1 << bz_shift
  int32_t cx;        // planes per workgroup along x
  int32_t by;        // block extent along y (threads per block = bz * by)
  // point sources folded into the E kernel (no periodic stage, no detector):
  // cell, and the waveform value to add at this step, in registration order
  int32_t npts;
  int64_t pt_cell[FOLDED_POINTS];
  const double* wave;  // waveform table, row s = source s
  int64_t wave_cap;
  int64_t q;
};

// ----------------------------------------------------------- device helpers
// The six one-sided differences of one cell, named after the curl terms:
// zy = difference of component z along y, etc. (grid.py:29-76).
template <typename T>
struct Diffs {
  T zy, yz, xz, zx, yx, xy;
};

// One CPML face at one cell (boundaries.py:409-487). A face normal to axis
// `ax` keeps psi for the two field components other than ax, C0 < C1
// (C0 = ax==0 ? 1 : 0, C1 = ax==2 ? 1 : 2). psi_q <- psi_q*b, plus
// (difference along ax of the other field's component 3-ax-C_q) * c where
// the slab index allows it (E side: l >= 1; H side: l <= t-2). The
// convolution term of C_q is +psi when ax follows C_q cyclically, -psi
// otherwise, so E[C0] gets +-(S inv)*psi_0 and E[C1] +-(S inv)*psi_1 (H with
// the opposite sign). Everything that depends on the axis is chosen with
// selects rather than branches, so no address is formed from values merged
// across control flow. Returns without touching v when the cell is outside
// the face's slab.
template <bool IS_E, typename T>
__device__ __forceinline__ void pml_face(const FaceDev& fd, const Dims& d, int i, int j, int k, T zx, T yx, T zy,
                                         T xy, T yz, T xz, T k0, T k1, T k2, T& v0, T& v1, T& v2) {
  const int ax = fd.axis;
  const int g = ax == 0 ? i : (ax == 1 ? j : k);
  const int n = ax == 0 ? d.nx : (ax == 1 ? d.ny : d.nz);
  const int l = fd.high ? g - (n - fd.t) : g;
  if (l < 0 || l >= fd.t) return;
  // slab: the grid with the normal axis shortened to t
  const int64_t sny = ax == 1 ? fd.t : d.ny;
  const int64_t snz = ax == 2 ? fd.t : d.nz;
  const int64_t li = ax == 0 ? l : i, lj = ax == 1 ? l : j, lk = ax == 2 ? l : k;
  const int64_t off = (li * sny + lj) * snz + lk;
  T* __restrict__ psi = reinterpret_cast<T*>(IS_E ? fd.psiE : fd.psiH);
  const double* bt = IS_E ? fd.bE : fd.bH;
  const double* ct = IS_E ? fd.cE : fd.cH;
  const T b = (T)bt[l], c = (T)ct[l];
  const T d0 = ax == 0 ? zx : (ax == 1 ? zy : yz);
  const T d1 = ax == 0 ? yx : (ax == 1 ? xy : xz);
  T p0 = psi[off] * b;
  T p1 = psi[fd.slab + off] * b;
  if (IS_E ? (l >= 1) : (l <= fd.t - 2)) {
    p0 = p0 + d0 * c;
    p1 = p1 + d1 * c;
  }
  psi[off] = p0;
  psi[fd.slab + off] = p1;
  const T r0 = (ax == 0 ? k1 : k0) * p0;
  const T r1 = (ax == 2 ? k1 : k2) * p1;
  T n0, n1, n2;
  if (IS_E) {
    // ax 0: E1 -= r0, E2 += r1; ax 1: E0 += r0, E2 -= r1; ax 2: E0 -= r0, E1 += r1
    n0 = ax == 1 ? v0 + r0 : (ax == 2 ? v0 - r0 : v0);
    n1 = ax == 0 ? v1 - r0 : (ax == 2 ? v1 + r1 : v1);
    n2 = ax == 0 ? v2 + r1 : (ax == 1 ? v2 - r1 : v2);
  } else {
    n0 = ax == 1 ? v0 - r0 : (ax == 2 ? v0 + r0 : v0);
    n1 = ax == 0 ? v1 + r0 : (ax == 2 ? v1 - r1 : v1);
    n2 = ax == 0 ? v2 - r1 : (ax == 1 ? v2 + r1 : v2);
  }
  v0 = n0;
  v1 = n1;
  v2 = n2;
}

// All faces in `faces` that contain cell (i,j,k), in registration order.
template <bool IS_E, typename T>
__device__ __forceinline__ void pml_cell(const FaceDev* __restrict__ faces, int nfaces, const Dims& d, int i, int j,
                                         int k, const Diffs<T>& df, const T kc[3], T v[3]) {
  for (int fi = 0; fi < nfaces; ++fi)
    pml_face<IS_E>(faces[fi], d, i, j, k, df.zx, df.yx, df.zy, df.xy, df.yz, df.xz, kc[0], kc[1], kc[2], v[0], v[1],
                   v[2]);
}

// Object updates of one cell, registration order (objects.py:118-269).
template <typename T>
__device__ __forceinline__ void objects_update_E(const ObjDev* __restrict__ objs, int nobjs, T sc, int i, int j,
                                                 int k, const T curl[3], T e[3]) {
  for (int o = 0; o < nobjs; ++o) {
    const ObjDev& ob = objs[o];
    if (i < ob.x0 || i >= ob.x1 || j < ob.y0 || j >= ob.y1 || k < ob.z0 || k >= ob.z1) continue;
    const int64_t lo = ((int64_t)(i - ob.x0) * (ob.y1 - ob.y0) + (j - ob.y0)) * (ob.z1 - ob.z0) + (k - ob.z0);
    const T* inv = reinterpret_cast<const T*>(ob.inv);
    if (ob.kind == FDTD_OBJ_PLAIN) {
#pragma unroll
      for (int c = 0; c < 3; ++c) e[c] = e[c] + (sc * inv[lo * 3 + c]) * curl[c];
    } else if (ob.kind == FDTD_OBJ_ABSORBING) {
      const T* fa = reinterpret_cast<const T*>(ob.absorb);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T f = fa[lo * 3 + c];
        e[c] = e[c] * ((T(1) - f) / (T(1) + f));
        e[c] = e[c] + ((sc * inv[lo * 3 + c]) * curl[c]) / (T(1) + f);
      }
    } else {
      const T* m = inv + lo * 9;
      T v[3];
#pragma unroll
      // the association numpy's einsum uses for this 3-term contraction
      for (int r = 0; r < 3; ++r) v[r] = (m[3 * r] * curl[0] + m[3 * r + 2] * curl[2]) + m[3 * r + 1] * curl[1];
#pragma unroll
      for (int r = 0; r < 3; ++r) e[r] = e[r] + sc * v[r];
    }
  }
}

// Loads and stores of the read-once streams (the updated field and the
// material); NT marks them non-temporal so the L2 keeps the other field,
// whose planes and rows are re-read as halos.
template <bool NT, typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void st_stream(T* p, T v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// Logical tile of hardware workgroup w: workgroups are dealt round-robin to
// the 8 XCDs (each with its own L2), so consecutive logical tiles -- which
// share halo planes -- are given to the same XCD.
__device__ __forceinline__ unsigned xcd_tile(unsigned w, unsigned nb) {
  const unsigned per = nb / 8, rem = nb % 8, x = w % 8, idx = w / 8;
  return x < rem ? x * (per + 1) + idx : rem * (per + 1) + (x - rem) * per + idx;
}

// ------------------------------------------------------------------ kernels
// Main E kernel. A workgroup owns a (y, z) tile of 256 columns (z fastest, one
// row of bz cells per wavefront when bz = 64) and marches along x over a chunk
// of `cx` planes: the x-neighbours come from registers of the previous plane,
// the z-neighbours from the neighbouring lane, the y-neighbours from cache.
template <typename T, bool OBJS, bool NT>
__global__ void __launch_bounds__(1024) k_main_E(MainArgs<T> a) {
  const Dims d = a.d;
  const int bzs = a.bz_shift, bz = 1 << bzs, by = a.by;
  const int ntz = (d.nz + bz - 1) >> bzs, nty = (d.ny + by - 1) / by;
  const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
  const int tz = tile % ntz, ty = (tile / ntz) % nty, tx = tile / (ntz * nty);
  const int lz = threadIdx.x & (bz - 1);
  const int k = tz * bz + lz, j = ty * by + (threadIdx.x >> bzs);
  const int i0 = tx * a.cx, i1 = min(i0 + a.cx, d.nx);
  const bool act = k < d.nz && j < d.ny;
  const int64_t sy = d.nz, sx = (int64_t)d.ny * d.nz, cs = d.cs;
  const T* __restrict__ Hx = a.G;
  const T* __restrict__ Hy = a.G + cs;
  const T* __restrict__ Hz = a.G + 2 * cs;
  T* __restrict__ F = a.F;
  const T* __restrict__ M = a.mat;
  const T zero = T(0);
  int64_t cell = (int64_t)i0 * sx + (int64_t)j * sy + k;
  T hz_prev = zero, hy_prev = zero;
  if (act && i0 > 0) {
    hz_prev = Hz[cell - sx];
    hy_prev = Hy[cell - sx];
  }
  for (int i = i0; i < i1; ++i, cell += sx) {
    T hx = zero, hy = zero, hz = zero;
    if (act) {
      hx = Hx[cell];
      hy = Hy[cell];
      hz = Hz[cell];
    }
    T hy_km = __shfl_up(hy, 1);
    T hx_km = __shfl_up(hx, 1);
    if (act) {
      if (lz == 0 && k > 0) {
        hy_km = Hy[cell - 1];
        hx_km = Hx[cell - 1];
      }
      T hz_jm = zero, hx_jm = zero;
      if (j > 0) {
        hz_jm = Hz[cell - sy];
        hx_jm = Hx[cell - sy];
      }
      // backward differences; an absent one contributes nothing (grid.py:67-74)
      Diffs<T> df;
      df.zy = j > 0 ? hz - hz_jm : zero;
      df.yz = k > 0 ? hy - hy_km : zero;
      df.xz = k > 0 ? hx - hx_km : zero;
      df.zx = i > 0 ? hz - hz_prev : zero;
      df.yx = i > 0 ? hy - hy_prev : zero;
      df.xy = j > 0 ? hx - hx_jm : zero;
      T curl[3];
      curl[0] = (zero + df.zy) - df.yz;
      curl[1] = (zero + df.xz) - df.zx;
      curl[2] = (zero + df.yx) - df.xy;
      T kc[3], e[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        kc[c] = a.sc * ld_stream<NT>(M + c * cs + cell);
        e[c] = ld_stream<NT>(F + c * cs + cell) + kc[c] * curl[c];
      }
      if (OBJS) objects_update_E(a.objs, a.nobjs, a.sc, i, j, k, curl, e);
      if (a.nfaces) pml_cell<true>(a.faces, a.nfaces, d, i, j, k, df, kc, e);
      for (int s = 0; s < a.npts; ++s)  // sources.py:109, registration order
        if (a.pt_cell[s] == cell) e[2] = e[2] + (T)a.wave[s * a.wave_cap + a.q];
#pragma unroll
      for (int c = 0; c < 3; ++c) st_stream<NT>(F + c * cs + cell, e[c]);
    }
    hz_prev = hz;
    hy_prev = hy;
  }
}

// Main H kernel: the mirror image of the E kernel. Its differences are
// forward ones, so workgroups take the tiles in decreasing order (the +y/+z
// neighbour tiles were read just before, by the same XCD) and march along -x,
// keeping E of plane i+1 in registers.
template <typename T, bool NT>
__global__ void __launch_bounds__(1024) k_main_H(MainArgs<T> a) {
  const Dims d = a.d;
  const int bzs = a.bz_shift, bz = 1 << bzs, by = a.by;
  const int ntz = (d.nz + bz - 1) >> bzs, nty = (d.ny + by - 1) / by;
  const unsigned tile = gridDim.x - 1 - xcd_tile(blockIdx.x, gridDim.x);
  const int tz = tile % ntz, ty = (tile / ntz) % nty, tx = tile / (ntz * nty);
  const int lz = threadIdx.x & (bz - 1);
  const int k = tz * bz + lz, j = ty * by + (threadIdx.x >> bzs);
  const int i0 = tx * a.cx, i1 = min(i0 + a.cx, d.nx);
  const bool act = k < d.nz && j < d.ny;
  const int64_t sy = d.nz, sx = (int64_t)d.ny * d.nz, cs = d.cs;
  const T* __restrict__ Ex = a.G;
  const T* __restrict__ Ey = a.G + cs;
  const T* __restrict__ Ez = a.G + 2 * cs;
  T* __restrict__ F = a.F;
  const T* __restrict__ M = a.mat;
  const T zero = T(0);
  int64_t cell = (int64_t)(i1 - 1) * sx + (int64_t)j * sy + k;
  T ez_next = zero, ey_next = zero;  // E at plane i+1
  if (act && i1 < d.nx) {
    ez_next = Ez[cell + sx];
    ey_next = Ey[cell + sx];
  }
  for (int i = i1 - 1; i >= i0; --i, cell -= sx) {
    T ex = zero, ey = zero, ez = zero;
    if (act) {
      ex = Ex[cell];
      ey = Ey[cell];
      ez = Ez[cell];
    }
    T ey_kp = __shfl_down(ey, 1);
    T ex_kp = __shfl_down(ex, 1);
    if (act) {
      if (lz == bz - 1 && k < d.nz - 1) {
        ey_kp = Ey[cell + 1];
        ex_kp = Ex[cell + 1];
      }
      T ez_jp = zero, ex_jp = zero;
      if (j < d.ny - 1) {
        ez_jp = Ez[cell + sy];
        ex_jp = Ex[cell + sy];
      }
      // forward differences; an absent one contributes nothing (grid.py:42-49)
      Diffs<T> df;
      df.zy = j < d.ny - 1 ? ez_jp - ez : zero;
      df.yz = k < d.nz - 1 ? ey_kp - ey : zero;
      df.xz = k < d.nz - 1 ? ex_kp - ex : zero;
      df.zx = i < d.nx - 1 ? ez_next - ez : zero;
      df.yx = i < d.nx - 1 ? ey_next - ey : zero;
      df.xy = j < d.ny - 1 ? ex_jp - ex : zero;
      T curl[3];
      curl[0] = (zero + df.zy) - df.yz;
      curl[1] = (zero + df.xz) - df.zx;
      curl[2] = (zero + df.yx) - df.xy;
      T kc[3], h[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        kc[c] = a.sc * ld_stream<NT>(M + c * cs + cell);
        h[c] = ld_stream<NT>(F + c * cs + cell) - kc[c] * curl[c];
      }
      // objects leave H unchanged (objects.py:131-137, 223-229, 271-277)
      if (a.nfaces) pml_cell<false>(a.faces, a.nfaces, d, i, j, k, df, kc, h);
#pragma unroll
      for (int c = 0; c < 3; ++c) st_stream<NT>(F + c * cs + cell, h[c]);
    }
    ez_next = ez;
    ey_next = ey;
  }
}

// Plain device copy of n 16-byte words: the bandwidth reference for the
// step kernels (a stream that reads and writes every byte once).
__global__ void __launch_bounds__(256) k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// One CPML face applied on its own (faces registered after a periodic
// boundary): F is the field being updated, G the other one. AX, the face's
// normal axis, is a template parameter so that no index depends on a runtime
// branch over the axis.
template <typename T, bool IS_E, int AX>
__global__ void __launch_bounds__(256) k_face(Dims d, T sc, T* F, const T* G, const T* mat, FaceDev fd) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= fd.slab) return;
  // slab-local coordinates, the slab laid out like the grid with axis AX of length t
  const int64_t snx = AX == 0 ? fd.t : d.nx, sny = AX == 1 ? fd.t : d.ny, snz = AX == 2 ? fd.t : d.nz;
  const int64_t li = p / (sny * snz), lj = (p / snz) % sny, lk = p % snz;
  (void)snx;
  const int l = (int)(AX == 0 ? li : (AX == 1 ? lj : lk));
  const int n = AX == 0 ? d.nx : (AX == 1 ? d.ny : d.nz);
  const int g = fd.high ? n - fd.t + l : l;
  const int i = AX == 0 ? g : (int)li, j = AX == 1 ? g : (int)lj, k = AX == 2 ? g : (int)lk;
  const int64_t cell = ((int64_t)i * d.ny + j) * d.nz + k;
  const int64_t st = AX == 0 ? (int64_t)d.ny * d.nz : (AX == 1 ? (int64_t)d.nz : 1);
  // the other field's components differenced by this face: 3-AX-C0, 3-AX-C1
  constexpr int S0 = 3 - AX - (AX == 0 ? 1 : 0), S1 = 3 - AX - (AX == 2 ? 1 : 2);
  const T* G0 = G + S0 * d.cs;
  const T* G1 = G + S1 * d.cs;
  T d0 = T(0), d1 = T(0);
  if (IS_E && l >= 1) {
    d0 = G0[cell] - G0[cell - st];
    d1 = G1[cell] - G1[cell - st];
  }
  if (!IS_E && l <= fd.t - 2) {
    d0 = G0[cell + st] - G0[cell];
    d1 = G1[cell + st] - G1[cell];
  }
  T kc[3], v[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    kc[c] = sc * mat[c * d.cs + cell];
    v[c] = F[c * d.cs + cell];
  }
  // pml_face reads the two normal-axis differences
  pml_face<IS_E>(fd, d, i, j, k, d0, d1, d0, d1, d0, d1, kc[0], kc[1], kc[2], v[0], v[1], v[2]);
#pragma unroll
  for (int c = 0; c < 3; ++c) F[c * d.cs + cell] = v[c];
}

// Periodic plane copy: to_low ? plane 0 <- plane n-1 : plane n-1 <- plane 0.
template <typename T>
__global__ void __launch_bounds__(256) k_periodic(Dims d, T* F, int axis, int to_low) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t n_plane, src, dst;
  if (axis == 0) {
    n_plane = (int64_t)d.ny * d.nz;
    if (p >= n_plane) return;
    const int64_t last = (int64_t)(d.nx - 1) * n_plane + p;
    src = to_low ? last : p;
    dst = to_low ? p : last;
  } else if (axis == 1) {
    n_plane = (int64_t)d.nx * d.nz;
    if (p >= n_plane) return;
    const int64_t i = p / d.nz, k = p % d.nz;
    const int64_t first = (i * d.ny) * d.nz + k, last = (i * d.ny + d.ny - 1) * d.nz + k;
    src = to_low ? last : first;
    dst = to_low ? first : last;
  } else {
    n_plane = (int64_t)d.nx * d.ny;
    if (p >= n_plane) return;
    const int64_t first = p * d.nz, last = p * d.nz + d.nz - 1;
    src = to_low ? last : first;
    dst = to_low ? first : last;
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) F[c * d.cs + dst] = F[c * d.cs + src];
}

template <typename T>
struct TailArgs {
  Dims d;
  T* F;  // field of this half step
  int32_t is_H;
  int32_t nsrc, ndet;
  const SrcDev* srcs;
  const DetDev* dets;
  const double* wave;
  int64_t wave_cap;
  int64_t q;
  int64_t counter;  // half steps of this kind already taken by the plan
};

// Sources in registration order, then detector samples; one workgroup so the
// ordering between sources touching the same cell and the detector reads is
// kept by workgroup barriers.
template <typename T>
__global__ void __launch_bounds__(256) k_tail(TailArgs<T> a) {
  const int tid = threadIdx.x;
  for (int s = 0; s < a.nsrc; ++s) {
    const SrcDev& src = a.srcs[s];
    if (a.is_H && src.kind != FDTD_SRC_PLANE) continue;
    const T w = (T)a.wave[s * a.wave_cap + a.q];
    T* Fc = a.F + (int64_t)(a.is_H ? src.compH : src.compE) * a.d.cs;
    for (int p = tid; p < src.n; p += blockDim.x) {
      const int64_t cell = src.cells[p];
      if (src.kind == FDTD_SRC_POINT)
        Fc[cell] = Fc[cell] + w;
      else if (src.kind == FDTD_SRC_LINE)
        Fc[cell] = Fc[cell] + (T)src.profile[p] * w;
      else
        Fc[cell] = (T)src.amplitude * w;
    }
    __syncthreads();
  }
  for (int t = 0; t < a.ndet; ++t) {
    const DetDev& det = a.dets[t];
    const int64_t slot = a.counter - (a.is_H ? det.slot0H : det.slot0E);
    T* buf = reinterpret_cast<T*>(a.is_H ? det.bufH : det.bufE) + slot * det.n * 3;
    for (int p = tid; p < det.n * 3; p += blockDim.x) {
      const int c = p % 3;
      buf[p] = a.F[c * a.d.cs + det.cells[p / 3]];
    }
  }
}

// ------------------------------------------------------------------- host
struct FaceHost {
  int axis, high, t;
  std::vector<double> bE, cE, bH, cH;
  void* psiE;
  void* psiH;
};
struct ObjHost {
  int kind, x0, x1, y0, y1, z0, z1;
  const void* inv;
  const void* absorb;
};
struct SrcHost {
  int kind, compE, compH;
  std::vector<int64_t> cells;
  std::vector<double> profile;
  double amplitude;
};
struct DetHost {
  std::vector<int64_t> cells;
  void* bufE;
  void* bufH;
  int64_t cap, countE, countH;
};
struct Bnd {
  int periodic;  // 1: periodic, 0: PML face
  int index;     // axis for periodic, face index for PML
};

size_t align_up(size_t v) { return (v + 255) & ~size_t(255); }

}  // namespace

struct FdtdPlan {
  int dtype = FDTD_F64;
  int device = 0;
  Dims d{};
  double sc = 0;
  void* E = nullptr;
  void* H = nullptr;
  const void* inv_eps = nullptr;
  const void* inv_mu = nullptr;
  std::vector<FaceHost> faces;
  std::vector<ObjHost> objs;
  std::vector<SrcHost> srcs;
  std::vector<DetHost> dets;
  std::vector<Bnd> order;
  const double* wave = nullptr;
  int64_t wave_cap = 0;
  bool finalized = false;
  // device tables
  void* dev = nullptr;
  FaceDev* d_faces = nullptr;  // faces in boundary registration order
  ObjDev* d_objs = nullptr;
  SrcDev* d_srcs = nullptr;
  DetDev* d_dets = nullptr;
  std::vector<FaceDev> h_faces;
  int n_fused = 0;           // leading faces folded into the main kernels
  std::vector<Bnd> stages;   // boundaries applied after the main kernel
  int64_t nE = 0, nH = 0;    // half steps taken (detector slot clocks)
  int cx = 1;                // x planes per workgroup of the main kernels
  int block_threads = 256;   // threads per workgroup of the main kernels
  bool fold_points = false;  // point sources added inside the E kernel
  bool nontemporal = true;   // read-once streams marked non-temporal
};

namespace {

int check_plan(const FdtdPlan* p) {
  if (!p) return fail(-1, "null plan");
  if (!p->finalized) return fail(-2, "plan not finalized");
  return 0;
}

int cell_ok(const FdtdPlan* p, int64_t c) { return c >= 0 && c < p->d.cs; }

void free_tables(FdtdPlan* p) {
  // kernels already queued may still read the old tables
  if (p->dev) (void)hipDeviceSynchronize();
  if (p->dev) (void)hipFree(p->dev);
  p->dev = nullptr;
  p->d_faces = nullptr;
  p->d_objs = nullptr;
  p->d_srcs = nullptr;
  p->d_dets = nullptr;
}

int select_device(const FdtdPlan* p) {
  int cur = -1;
  FDTD_HIP(hipGetDevice(&cur));
  if (cur != p->device) FDTD_HIP(hipSetDevice(p->device));
  return 0;
}

template <typename T>
MainArgs<T> main_args(const FdtdPlan* p, bool is_H) {
  MainArgs<T> a{};
  a.d = p->d;
  a.sc = (T)p->sc;
  a.F = reinterpret_cast<T*>(is_H ? p->H : p->E);
  a.G = reinterpret_cast<const T*>(is_H ? p->E : p->H);
  a.mat = reinterpret_cast<const T*>(is_H ? p->inv_mu : p->inv_eps);
  a.faces = p->d_faces;
  a.nfaces = p->n_fused;
  a.objs = p->d_objs;
  a.nobjs = is_H ? 0 : (int)p->objs.size();
  int bzs = 0;
  while ((1 << bzs) < p->d.nz && bzs < 6) ++bzs;
  a.bz_shift = bzs;
  a.cx = p->cx;
  a.by = std::max(1, p->block_threads >> bzs);
  a.npts = 0;
  return a;
}

template <typename T>
void launch_main(const FdtdPlan* p, const MainArgs<T>& a, bool is_H, hipStream_t st) {
  const int bz = 1 << a.bz_shift, by = a.by;
  const unsigned tiles = (unsigned)(((p->d.nz + bz - 1) / bz) * ((p->d.ny + by - 1) / by));
  const dim3 grid(tiles * (unsigned)((p->d.nx + a.cx - 1) / a.cx)), block(bz * by);
  if (p->nontemporal) {
    if (is_H)
      hipLaunchKernelGGL((k_main_H<T, true>), grid, block, 0, st, a);
    else if (a.nobjs)
      hipLaunchKernelGGL((k_main_E<T, true, true>), grid, block, 0, st, a);
    else
      hipLaunchKernelGGL((k_main_E<T, false, true>), grid, block, 0, st, a);
  } else {
    if (is_H)
      hipLaunchKernelGGL((k_main_H<T, false>), grid, block, 0, st, a);
    else if (a.nobjs)
      hipLaunchKernelGGL((k_main_E<T, true, false>), grid, block, 0, st, a);
    else
      hipLaunchKernelGGL((k_main_E<T, false, false>), grid, block, 0, st, a);
  }
}

template <typename T>
int launch_half(FdtdPlan* p, bool is_H, int64_t q, hipStream_t st) {
  MainArgs<T> a = main_args<T>(p, is_H);
  if (!is_H && p->fold_points) {
    a.npts = (int32_t)p->srcs.size();
    for (int s = 0; s < a.npts; ++s) a.pt_cell[s] = p->srcs[s].cells[0];
    a.wave = p->wave;
    a.wave_cap = p->wave_cap;
    a.q = q;
  }
  launch_main<T>(p, a, is_H, st);
  for (const Bnd& b : p->stages) {
    if (b.periodic) {
      const int axis = b.index;
      const int n_axis = axis == 0 ? p->d.nx : (axis == 1 ? p->d.ny : p->d.nz);
      if (n_axis <= 1) continue;  // plane 0 is plane -1: the copy is the identity
      const int64_t n_plane = p->d.cs / n_axis;
      hipLaunchKernelGGL(k_periodic<T>, dim3((unsigned)((n_plane + 255) / 256)), dim3(256), 0, st, p->d, a.F,
                         axis, is_H ? 0 : 1);
    } else {
      const FaceDev& fd = p->h_faces[b.index];
      const dim3 nb((unsigned)((fd.slab + 255) / 256)), nt(256);
      switch (fd.axis * 2 + (is_H ? 1 : 0)) {
        case 0: hipLaunchKernelGGL((k_face<T, true, 0>), nb, nt, 0, st, p->d, a.sc, a.F, a.G, a.mat, fd); break;
        case 1: hipLaunchKernelGGL((k_face<T, false, 0>), nb, nt, 0, st, p->d, a.sc, a.F, a.G, a.mat, fd); break;
        case 2: hipLaunchKernelGGL((k_face<T, true, 1>), nb, nt, 0, st, p->d, a.sc, a.F, a.G, a.mat, fd); break;
        case 3: hipLaunchKernelGGL((k_face<T, false, 1>), nb, nt, 0, st, p->d, a.sc, a.F, a.G, a.mat, fd); break;
        case 4: hipLaunchKernelGGL((k_face<T, true, 2>), nb, nt, 0, st, p->d, a.sc, a.F, a.G, a.mat, fd); break;
        default: hipLaunchKernelGGL((k_face<T, false, 2>), nb, nt, 0, st, p->d, a.sc, a.F, a.G, a.mat, fd); break;
      }
    }
  }
  bool any_src = false;
  if (!p->fold_points)
    for (const SrcHost& s : p->srcs) any_src |= (!is_H || s.kind == FDTD_SRC_PLANE);
  if (any_src || !p->dets.empty()) {
    TailArgs<T> t{};
    t.d = p->d;
    t.F = a.F;
    t.is_H = is_H ? 1 : 0;
    t.nsrc = (int)p->srcs.size();
    t.ndet = (int)p->dets.size();
    t.srcs = p->d_srcs;
    t.dets = p->d_dets;
    t.wave = p->wave;
    t.wave_cap = p->wave_cap;
    t.q = q;
    t.counter = is_H ? p->nH : p->nE;
    hipLaunchKernelGGL(k_tail<T>, dim3(1), dim3(256), 0, st, t);
  }
  FDTD_HIP(hipGetLastError());
  if (is_H) {
    ++p->nH;
    for (DetHost& dh : p->dets) ++dh.countH;
  } else {
    ++p->nE;
    for (DetHost& dh : p->dets) ++dh.countE;
  }
  return 0;
}

int check_step_room(const FdtdPlan* p, int64_t q0, int64_t nsteps, bool is_H_only, bool e_only) {
  if (q0 < 0) return fail(-1, "negative time index");
  if (!p->srcs.empty()) {
    if (!p->wave || p->wave_cap < q0 + nsteps)
      return fail(-4, "waveform table too short for time index " + std::to_string(q0 + nsteps - 1));
  }
  for (size_t i = 0; i < p->dets.size(); ++i) {
    const DetHost& dh = p->dets[i];
    const int64_t needE = dh.countE + (is_H_only ? 0 : nsteps);
    const int64_t needH = dh.countH + (e_only ? 0 : nsteps);
    if (needE > dh.cap || needH > dh.cap)
      return fail(-4, "detector " + std::to_string(i) + " buffer too small");
  }
  return 0;
}

int half_dispatch(FdtdPlan* p, bool is_H, int64_t q, hipStream_t st) {
  return p->dtype == FDTD_F64 ? launch_half<double>(p, is_H, q, st) : launch_half<float>(p, is_H, q, st);
}

}  // namespace

// ------------------------------------------------------------------- C ABI
extern "C" {

const char* fdtd_version(void) { return "fdtd_hip 0.1 (gfx950)"; }

const char* fdtd_last_error(void) { return g_err.c_str(); }

int fdtd_plan_create(int dtype, int nx, int ny, int nz, double courant_number, void* E, void* H,
                     const void* inv_eps, const void* inv_mu, FdtdPlan** out) {
  if (!out) return fail(-1, "null output pointer");
  *out = nullptr;
  if (dtype != FDTD_F64 && dtype != FDTD_F32) return fail(-1, "unknown dtype");
  if (nx < 1 || ny < 1 || nz < 1) return fail(-1, "grid dimensions must be positive");
  if (nx > 65535) return fail(-1, "nx above 65535 is not supported");
  if (!E || !H || !inv_eps || !inv_mu) return fail(-1, "null field or material pointer");
  FdtdPlan* p = new FdtdPlan();
  p->dtype = dtype;
  p->d.nx = nx;
  p->d.ny = ny;
  p->d.nz = nz;
  p->d.cs = (int64_t)nx * ny * nz;
  p->sc = courant_number;
  p->E = E;
  p->H = H;
  p->inv_eps = inv_eps;
  p->inv_mu = inv_mu;
  if (hipGetDevice(&p->device) != hipSuccess) {
    delete p;
    return fail(-3, "hipGetDevice failed");
  }
  *out = p;
  return 0;
}

int fdtd_plan_destroy(FdtdPlan* plan) {
  if (!plan) return 0;
  free_tables(plan);
  delete plan;
  return 0;
}

int fdtd_plan_add_pml(FdtdPlan* p, int axis, int high, int thickness, const double* bE, const double* cE,
                      const double* bH, const double* cH, void* psiE, void* psiH) {
  if (!p) return fail(-1, "null plan");
  if (axis < 0 || axis > 2) return fail(-1, "PML axis must be 0, 1 or 2");
  const int n = axis == 0 ? p->d.nx : (axis == 1 ? p->d.ny : p->d.nz);
  if (thickness < 1 || thickness > n) return fail(-1, "PML thickness out of range");
  if (!bE || !cE || !bH || !cH || !psiE || !psiH) return fail(-1, "null PML array");
  FaceHost f;
  f.axis = axis;
  f.high = high ? 1 : 0;
  f.t = thickness;
  f.bE.assign(bE, bE + thickness);
  f.cE.assign(cE, cE + thickness);
  f.bH.assign(bH, bH + thickness);
  f.cH.assign(cH, cH + thickness);
  f.psiE = psiE;
  f.psiH = psiH;
  p->order.push_back({0, (int)p->faces.size()});
  p->faces.push_back(std::move(f));
  p->finalized = false;
  return 0;
}

int fdtd_plan_add_periodic(FdtdPlan* p, int axis) {
  if (!p) return fail(-1, "null plan");
  if (axis < 0 || axis > 2) return fail(-1, "periodic axis must be 0, 1 or 2");
  p->order.push_back({1, axis});
  p->finalized = false;
  return 0;
}

int fdtd_plan_add_object(FdtdPlan* p, int kind, int x0, int x1, int y0, int y1, int z0, int z1, const void* inv,
                         const void* absorption) {
  if (!p) return fail(-1, "null plan");
  if (kind < FDTD_OBJ_PLAIN || kind > FDTD_OBJ_ANISOTROPIC) return fail(-1, "unknown object kind");
  if (x0 < 0 || y0 < 0 || z0 < 0 || x1 > p->d.nx || y1 > p->d.ny || z1 > p->d.nz || x0 >= x1 || y0 >= y1 ||
      z0 >= z1)
    return fail(-1, "object box outside the grid or empty");
  if (!inv) return fail(-1, "null object permittivity");
  if (kind == FDTD_OBJ_ABSORBING && !absorption) return fail(-1, "null absorption factor");
  p->objs.push_back({kind, x0, x1, y0, y1, z0, z1, inv, absorption});
  p->finalized = false;
  return 0;
}

int fdtd_plan_add_source(FdtdPlan* p, int kind, int n, const int64_t* cells, const double* profile, double amplitude,
                         int comp_E, int comp_H) {
  if (!p) return fail(-1, "null plan");
  if (kind < FDTD_SRC_POINT || kind > FDTD_SRC_PLANE) return fail(-1, "unknown source kind");
  if (n < 1 || !cells) return fail(-1, "source without cells");
  if (comp_E < 0 || comp_E > 2 || (kind == FDTD_SRC_PLANE && (comp_H < 0 || comp_H > 2)))
    return fail(-1, "source component out of range");
  if (kind == FDTD_SRC_LINE && !profile) return fail(-1, "line source without profile");
  SrcHost s;
  s.kind = kind;
  s.compE = comp_E;
  s.compH = kind == FDTD_SRC_PLANE ? comp_H : 0;
  s.cells.assign(cells, cells + n);
  for (int64_t c : s.cells)
    if (!cell_ok(p, c)) return fail(-1, "source cell outside the grid");
  if (profile) s.profile.assign(profile, profile + n);
  s.amplitude = amplitude;
  p->srcs.push_back(std::move(s));
  p->finalized = false;
  return 0;
}

int fdtd_plan_set_waveforms(FdtdPlan* p, const double* table, int64_t capacity) {
  if (!p) return fail(-1, "null plan");
  if (capacity < 0 || (capacity > 0 && !table)) return fail(-1, "bad waveform table");
  p->wave = table;
  p->wave_cap = capacity;
  return 0;
}

int fdtd_plan_add_detector(FdtdPlan* p, int n, const int64_t* cells, void* bufE, void* bufH, int64_t capacity,
                           int64_t count_E, int64_t count_H) {
  if (!p) return fail(-1, "null plan");
  if (n < 1 || !cells) return fail(-1, "detector without cells");
  if (!bufE || !bufH || capacity < 0 || count_E < 0 || count_H < 0 || count_E > capacity || count_H > capacity)
    return fail(-1, "bad detector buffers");
  DetHost dh;
  dh.cells.assign(cells, cells + n);
  for (int64_t c : dh.cells)
    if (!cell_ok(p, c)) return fail(-1, "detector cell outside the grid");
  dh.bufE = bufE;
  dh.bufH = bufH;
  dh.cap = capacity;
  dh.countE = count_E;
  dh.countH = count_H;
  p->dets.push_back(std::move(dh));
  p->finalized = false;
  return 0;
}

int fdtd_plan_finalize(FdtdPlan* p) {
  if (!p) return fail(-1, "null plan");
  if (int rc = select_device(p)) return rc;
  free_tables(p);
  // boundaries before the first periodic one fold into the main kernels
  p->n_fused = 0;
  p->stages.clear();
  bool split = false;
  for (const Bnd& b : p->order) {
    if (b.periodic) split = true;
    if (!split && !b.periodic)
      ++p->n_fused;
    else
      p->stages.push_back(b);
  }
  // launch shape of the main kernels: workgroup size, then the x-chunk --
  // enough workgroups to fill 256 CUs many times over, long enough chunks to
  // amortise the halo plane
  {
    if (const char* env = getenv("FDTD_HIP_BLOCK")) {
      const int bt = atoi(env);
      if (bt == 256 || bt == 512 || bt == 1024) p->block_threads = bt;
    }
    int bzs = 0;
    while ((1 << bzs) < p->d.nz && bzs < 6) ++bzs;
    const int bz = 1 << bzs, by = std::max(1, p->block_threads >> bzs);
    const int64_t tiles = (int64_t)((p->d.nz + bz - 1) / bz) * ((p->d.ny + by - 1) / by);
    int64_t cx = tiles * p->d.nx / 32768;
    if (const char* env = getenv("FDTD_HIP_CX")) cx = atoi(env);
    if (const char* env = getenv("FDTD_HIP_NT")) p->nontemporal = atoi(env) != 0;
    p->cx = (int)std::max<int64_t>(1, std::min<int64_t>(cx, p->d.nx));
  }
  // point sources fold into the E kernel when nothing runs between the main
  // kernel and the sources (no periodic or staged face) and no detector samples
  p->fold_points = !p->srcs.empty() && (int)p->srcs.size() <= FOLDED_POINTS && p->dets.empty();
  for (const SrcHost& s : p->srcs) p->fold_points = p->fold_points && s.kind == FDTD_SRC_POINT;
  for (const Bnd& b : p->order) p->fold_points = p->fold_points && !b.periodic;
  // byte layout of the table block
  const size_t nf = p->faces.size(), no = p->objs.size(), ns = p->srcs.size(), nd = p->dets.size();
  size_t off = 0;
  const size_t o_faces = off;
  off = align_up(off + nf * sizeof(FaceDev));
  const size_t o_objs = off;
  off = align_up(off + no * sizeof(ObjDev));
  const size_t o_srcs = off;
  off = align_up(off + ns * sizeof(SrcDev));
  const size_t o_dets = off;
  off = align_up(off + nd * sizeof(DetDev));
  std::vector<size_t> o_coef(nf), o_scell(ns), o_sprof(ns), o_dcell(nd);
  for (size_t i = 0; i < nf; ++i) {
    o_coef[i] = off;
    off = align_up(off + 4 * p->faces[i].t * sizeof(double));
  }
  for (size_t i = 0; i < ns; ++i) {
    o_scell[i] = off;
    off = align_up(off + p->srcs[i].cells.size() * sizeof(int64_t));
    o_sprof[i] = off;
    off = align_up(off + p->srcs[i].profile.size() * sizeof(double));
  }
  for (size_t i = 0; i < nd; ++i) {
    o_dcell[i] = off;
    off = align_up(off + p->dets[i].cells.size() * sizeof(int64_t));
  }
  const size_t total = off > 0 ? off : 256;
  std::vector<unsigned char> host(total, 0);
  char* base = nullptr;
  FDTD_HIP(hipMalloc(&p->dev, total));
  base = reinterpret_cast<char*>(p->dev);
  p->h_faces.clear();
  for (size_t i = 0; i < nf; ++i) {
    const FaceHost& f = p->faces[i];
    double* coef = reinterpret_cast<double*>(host.data() + o_coef[i]);
    std::memcpy(coef, f.bE.data(), f.t * sizeof(double));
    std::memcpy(coef + f.t, f.cE.data(), f.t * sizeof(double));
    std::memcpy(coef + 2 * f.t, f.bH.data(), f.t * sizeof(double));
    std::memcpy(coef + 3 * f.t, f.cH.data(), f.t * sizeof(double));
    const double* dcoef = reinterpret_cast<const double*>(base + o_coef[i]);
    FaceDev fd{};
    fd.axis = f.axis;
    fd.high = f.high;
    fd.t = f.t;
    fd.bE = dcoef;
    fd.cE = dcoef + f.t;
    fd.bH = dcoef + 2 * f.t;
    fd.cH = dcoef + 3 * f.t;
    fd.psiE = f.psiE;
    fd.psiH = f.psiH;
    fd.slab = p->d.cs / (f.axis == 0 ? p->d.nx : (f.axis == 1 ? p->d.ny : p->d.nz)) * f.t;
    p->h_faces.push_back(fd);
  }
  // fused faces first (registration order), then the staged ones
  std::vector<FaceDev> ordered;
  for (const Bnd& b : p->order)
    if (!b.periodic) ordered.push_back(p->h_faces[b.index]);
  // faces are added in registration order, so `ordered` equals h_faces
  std::memcpy(host.data() + o_faces, ordered.data(), nf * sizeof(FaceDev));
  for (size_t i = 0; i < no; ++i) {
    const ObjHost& o = p->objs[i];
    ObjDev od{o.kind, o.x0, o.x1, o.y0, o.y1, o.z0, o.z1, 0, o.inv, o.absorb};
    std::memcpy(host.data() + o_objs + i * sizeof(ObjDev), &od, sizeof(ObjDev));
  }
  for (size_t i = 0; i < ns; ++i) {
    const SrcHost& s = p->srcs[i];
    std::memcpy(host.data() + o_scell[i], s.cells.data(), s.cells.size() * sizeof(int64_t));
    if (!s.profile.empty())
      std::memcpy(host.data() + o_sprof[i], s.profile.data(), s.profile.size() * sizeof(double));
    SrcDev sd{};
    sd.kind = s.kind;
    sd.n = (int32_t)s.cells.size();
    sd.compE = s.compE;
    sd.compH = s.compH;
    sd.cells = reinterpret_cast<const int64_t*>(base + o_scell[i]);
    sd.profile = s.profile.empty() ? nullptr : reinterpret_cast<const double*>(base + o_sprof[i]);
    sd.amplitude = s.amplitude;
    std::memcpy(host.data() + o_srcs + i * sizeof(SrcDev), &sd, sizeof(SrcDev));
  }
  for (size_t i = 0; i < nd; ++i) {
    const DetHost& dh = p->dets[i];
    std::memcpy(host.data() + o_dcell[i], dh.cells.data(), dh.cells.size() * sizeof(int64_t));
    DetDev dd{};
    dd.n = (int32_t)dh.cells.size();
    dd.cells = reinterpret_cast<const int64_t*>(base + o_dcell[i]);
    dd.bufE = dh.bufE;
    dd.bufH = dh.bufH;
    dd.slot0E = p->nE - dh.countE;
    dd.slot0H = p->nH - dh.countH;
    std::memcpy(host.data() + o_dets + i * sizeof(DetDev), &dd, sizeof(DetDev));
  }
  FDTD_HIP(hipMemcpy(p->dev, host.data(), total, hipMemcpyHostToDevice));
  p->d_faces = reinterpret_cast<FaceDev*>(base + o_faces);
  p->d_objs = reinterpret_cast<ObjDev*>(base + o_objs);
  p->d_srcs = reinterpret_cast<SrcDev*>(base + o_srcs);
  p->d_dets = reinterpret_cast<DetDev*>(base + o_dets);
  p->finalized = true;
  return 0;
}

int fdtd_plan_detector_counts(const FdtdPlan* p, int index, int64_t* count_E, int64_t* count_H) {
  if (!p) return fail(-1, "null plan");
  if (index < 0 || index >= (int)p->dets.size()) return fail(-1, "detector index out of range");
  if (count_E) *count_E = p->dets[index].countE;
  if (count_H) *count_H = p->dets[index].countH;
  return 0;
}

int fdtd_half_step_E(FdtdPlan* p, int64_t q, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (int rc = check_step_room(p, q, 1, false, true)) return rc;
  if (int rc = select_device(p)) return rc;
  return half_dispatch(p, false, q, reinterpret_cast<hipStream_t>(stream));
}

int fdtd_half_step_H(FdtdPlan* p, int64_t q, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (int rc = check_step_room(p, q, 1, true, false)) return rc;
  if (int rc = select_device(p)) return rc;
  return half_dispatch(p, true, q, reinterpret_cast<hipStream_t>(stream));
}

int fdtd_run(FdtdPlan* p, int64_t q0, int64_t nsteps, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (nsteps < 0) return fail(-1, "negative step count");
  if (int rc = check_step_room(p, q0, nsteps, false, false)) return rc;
  if (int rc = select_device(p)) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  for (int64_t s = 0; s < nsteps; ++s) {
    if (int rc = half_dispatch(p, false, q0 + s, st)) return rc;
    if (int rc = half_dispatch(p, true, q0 + s, st)) return rc;
  }
  return 0;
}

int fdtd_bench_main_kernel(FdtdPlan* p, int64_t reps, void* stream, float* ms_E, float* ms_H) {
  if (int rc = check_plan(p)) return rc;
  if (reps < 1 || reps > 4096 || !ms_E || !ms_H) return fail(-1, "bad benchmark arguments");
  if (int rc = select_device(p)) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  std::vector<hipEvent_t> ev(2 * reps + 1);
  for (auto& e : ev) FDTD_HIP(hipEventCreate(&e));
  FDTD_HIP(hipEventRecord(ev[0], st));
  for (int64_t r = 0; r < reps; ++r) {
    for (int h = 0; h < 2; ++h) {
      if (p->dtype == FDTD_F64)
        launch_main<double>(p, main_args<double>(p, h != 0), h != 0, st);
      else
        launch_main<float>(p, main_args<float>(p, h != 0), h != 0, st);
      FDTD_HIP(hipEventRecord(ev[2 * r + h + 1], st));
    }
  }
  FDTD_HIP(hipGetLastError());
  FDTD_HIP(hipEventSynchronize(ev[2 * reps]));
  double tE = 0, tH = 0;
  for (int64_t r = 0; r < reps; ++r) {
    float a = 0, b = 0;
    FDTD_HIP(hipEventElapsedTime(&a, ev[2 * r], ev[2 * r + 1]));
    FDTD_HIP(hipEventElapsedTime(&b, ev[2 * r + 1], ev[2 * r + 2]));
    tE += a;
    tH += b;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  *ms_E = (float)(tE / reps);
  *ms_H = (float)(tH / reps);
  return 0;
}

int fdtd_bench_copy(void* dst, const void* src, int64_t bytes, int64_t reps, void* stream, float* ms) {
  if (!dst || !src || bytes < 16 || (bytes & 15) || reps < 1 || !ms) return fail(-1, "bad copy benchmark arguments");
  if ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15)
    return fail(-1, "copy buffers must be 16-byte aligned");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  int dev = 0, cus = 0;
  FDTD_HIP(hipGetDevice(&dev));
  FDTD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int64_t n = bytes / 16;
  // 4 workgroups per CU: the fastest shape measured for a plain streaming copy
  const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, (int64_t)cus * 4);
  hipEvent_t e0, e1;
  FDTD_HIP(hipEventCreate(&e0));
  FDTD_HIP(hipEventCreate(&e1));
  FDTD_HIP(hipEventRecord(e0, st));
  for (int64_t r = 0; r < reps; ++r)
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, st, reinterpret_cast<const uint4*>(src),
                       reinterpret_cast<uint4*>(dst), n);
  FDTD_HIP(hipGetLastError());
  FDTD_HIP(hipEventRecord(e1, st));
  FDTD_HIP(hipEventSynchronize(e1));
  FDTD_HIP(hipEventElapsedTime(ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 0;
}

}  // extern "C"
