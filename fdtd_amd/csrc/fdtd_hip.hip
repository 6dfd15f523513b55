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
//   k_main_E / k_main_H  one thread per cell: curl, material update, objects
//                        and every CPML face registered before the first
//                        periodic boundary. Memory-bound (~0.16 flop/B).
//   k_face_E / k_face_H  one CPML face over its slab, for faces registered
//                        after a periodic boundary (order matters there).
//   k_periodic           plane copy of a periodic boundary.
//   k_tail               one workgroup: sources (in order) then detectors.

#include <hip/hip_runtime.h>

#include <cstdint>
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
};

// ----------------------------------------------------------- device helpers
__device__ __forceinline__ int face_local(const FaceDev& f, const Dims& d, int i, int j, int k) {
  const int g = f.axis == 0 ? i : (f.axis == 1 ? j : k);
  const int n = f.axis == 0 ? d.nx : (f.axis == 1 ? d.ny : d.nz);
  return f.high ? g - (n - f.t) : g;
}

__device__ __forceinline__ int64_t face_offset(const FaceDev& f, const Dims& d, int i, int j, int k, int l) {
  if (f.axis == 0) return ((int64_t)l * d.ny + j) * d.nz + k;
  if (f.axis == 1) return ((int64_t)i * f.t + l) * d.nz + k;
  return ((int64_t)i * d.ny + j) * f.t + l;
}

__device__ __forceinline__ int64_t axis_stride(const Dims& d, int axis) {
  return axis == 0 ? (int64_t)d.ny * d.nz : (axis == 1 ? (int64_t)d.nz : 1);
}

// The CPML face's two live components: the field components other than the
// normal axis, in increasing order. The convolution term of component c is
// +psi when the axis follows c cyclically, -psi otherwise
// (boundaries.py:457-459, 485-487 with only the normal-axis psi non-zero).
__device__ __forceinline__ int face_comp(int axis, int q) {
  return q == 0 ? (axis == 0 ? 1 : 0) : (axis == 2 ? 1 : 2);
}

// E-side psi update of one face cell plus the E correction
// (boundaries.py:433-459 then :409-419). `kc` holds S*inv_eps per component.
template <typename T>
__device__ __forceinline__ void face_update_E(const FaceDev& f, const Dims& d, const T* __restrict__ H,
                                              int64_t cell, int l, int64_t off, const T kc[3], T e[3]) {
  T* psi = reinterpret_cast<T*>(f.psiE);
  const T b = (T)f.bE[l];
  const T c = (T)f.cE[l];
  const int64_t st = axis_stride(d, f.axis);
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int comp = face_comp(f.axis, q);
    const T* Hs = H + (int64_t)(3 - f.axis - comp) * d.cs;
    T p = psi[q * f.slab + off] * b;
    if (l >= 1) {
      const T dh = Hs[cell] - Hs[cell - st];
      p = p + dh * c;
    }
    psi[q * f.slab + off] = p;
    const T corr = kc[comp] * p;
    e[comp] = (f.axis == (comp + 1) % 3) ? e[comp] + corr : e[comp] - corr;
  }
}

// H-side psi update plus the H correction (boundaries.py:461-487, :421-431).
template <typename T>
__device__ __forceinline__ void face_update_H(const FaceDev& f, const Dims& d, const T* __restrict__ E,
                                              int64_t cell, int l, int64_t off, const T kc[3], T h[3]) {
  T* psi = reinterpret_cast<T*>(f.psiH);
  const T b = (T)f.bH[l];
  const T c = (T)f.cH[l];
  const int64_t st = axis_stride(d, f.axis);
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int comp = face_comp(f.axis, q);
    const T* Es = E + (int64_t)(3 - f.axis - comp) * d.cs;
    T p = psi[q * f.slab + off] * b;
    if (l <= f.t - 2) {
      const T de = Es[cell + st] - Es[cell];
      p = p + de * c;
    }
    psi[q * f.slab + off] = p;
    const T corr = kc[comp] * p;
    h[comp] = (f.axis == (comp + 1) % 3) ? h[comp] - corr : h[comp] + corr;
  }
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

// ------------------------------------------------------------------ kernels
template <typename T>
__global__ void __launch_bounds__(256) k_main_E(MainArgs<T> a) {
  const int bz = 1 << a.bz_shift;
  const int k = blockIdx.x * bz + (threadIdx.x & (bz - 1));
  const int j = blockIdx.y * (256 >> a.bz_shift) + (threadIdx.x >> a.bz_shift);
  const int i = blockIdx.z;
  const Dims d = a.d;
  if (k >= d.nz || j >= d.ny) return;
  const int64_t sy = d.nz, sx = (int64_t)d.ny * d.nz, cs = d.cs;
  const int64_t cell = i * sx + j * sy + k;
  const T* __restrict__ Hx = a.G;
  const T* __restrict__ Hy = a.G + cs;
  const T* __restrict__ Hz = a.G + 2 * cs;
  const T hx = Hx[cell], hy = Hy[cell], hz = Hz[cell];
  const T zero = T(0);
  // backward differences; an absent one contributes nothing (grid.py:67-74)
  const T dHz_y = j > 0 ? hz - Hz[cell - sy] : zero;
  const T dHy_z = k > 0 ? hy - Hy[cell - 1] : zero;
  const T dHx_z = k > 0 ? hx - Hx[cell - 1] : zero;
  const T dHz_x = i > 0 ? hz - Hz[cell - sx] : zero;
  const T dHy_x = i > 0 ? hy - Hy[cell - sx] : zero;
  const T dHx_y = j > 0 ? hx - Hx[cell - sy] : zero;
  T curl[3];
  curl[0] = (zero + dHz_y) - dHy_z;
  curl[1] = (zero + dHx_z) - dHz_x;
  curl[2] = (zero + dHy_x) - dHx_y;
  T kc[3], e[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    kc[c] = a.sc * a.mat[c * cs + cell];
    e[c] = a.F[c * cs + cell] + kc[c] * curl[c];
  }
  if (a.nobjs) objects_update_E(a.objs, a.nobjs, a.sc, i, j, k, curl, e);
  for (int f = 0; f < a.nfaces; ++f) {
    const FaceDev& fd = a.faces[f];
    const int l = face_local(fd, d, i, j, k);
    if (l < 0 || l >= fd.t) continue;
    face_update_E(fd, d, a.G, cell, l, face_offset(fd, d, i, j, k, l), kc, e);
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) a.F[c * cs + cell] = e[c];
}

template <typename T>
__global__ void __launch_bounds__(256) k_main_H(MainArgs<T> a) {
  const int bz = 1 << a.bz_shift;
  const int k = blockIdx.x * bz + (threadIdx.x & (bz - 1));
  const int j = blockIdx.y * (256 >> a.bz_shift) + (threadIdx.x >> a.bz_shift);
  const int i = blockIdx.z;
  const Dims d = a.d;
  if (k >= d.nz || j >= d.ny) return;
  const int64_t sy = d.nz, sx = (int64_t)d.ny * d.nz, cs = d.cs;
  const int64_t cell = i * sx + j * sy + k;
  const T* __restrict__ Ex = a.G;
  const T* __restrict__ Ey = a.G + cs;
  const T* __restrict__ Ez = a.G + 2 * cs;
  const T ex = Ex[cell], ey = Ey[cell], ez = Ez[cell];
  const T zero = T(0);
  // forward differences; an absent one contributes nothing (grid.py:42-49)
  const T dEz_y = j < d.ny - 1 ? Ez[cell + sy] - ez : zero;
  const T dEy_z = k < d.nz - 1 ? Ey[cell + 1] - ey : zero;
  const T dEx_z = k < d.nz - 1 ? Ex[cell + 1] - ex : zero;
  const T dEz_x = i < d.nx - 1 ? Ez[cell + sx] - ez : zero;
  const T dEy_x = i < d.nx - 1 ? Ey[cell + sx] - ey : zero;
  const T dEx_y = j < d.ny - 1 ? Ex[cell + sy] - ex : zero;
  T curl[3];
  curl[0] = (zero + dEz_y) - dEy_z;
  curl[1] = (zero + dEx_z) - dEz_x;
  curl[2] = (zero + dEy_x) - dEx_y;
  T kc[3], h[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    kc[c] = a.sc * a.mat[c * cs + cell];
    h[c] = a.F[c * cs + cell] - kc[c] * curl[c];
  }
  // objects leave H unchanged (objects.py:131-137, 223-229, 271-277)
  for (int f = 0; f < a.nfaces; ++f) {
    const FaceDev& fd = a.faces[f];
    const int l = face_local(fd, d, i, j, k);
    if (l < 0 || l >= fd.t) continue;
    face_update_H(fd, d, a.G, cell, l, face_offset(fd, d, i, j, k, l), kc, h);
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) a.F[c * cs + cell] = h[c];
}

// One CPML face applied on its own (faces registered after a periodic
// boundary): F is the field being updated, G the other one.
template <typename T, bool IS_E>
__global__ void __launch_bounds__(256) k_face(Dims d, T sc, T* F, const T* G, const T* mat, FaceDev fd) {
  const int64_t n = fd.slab;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int i, j, k, l;
  if (fd.axis == 0) {
    k = p % d.nz; j = (p / d.nz) % d.ny; l = p / ((int64_t)d.nz * d.ny);
    i = fd.high ? d.nx - fd.t + l : l;
  } else if (fd.axis == 1) {
    k = p % d.nz; l = (p / d.nz) % fd.t; i = p / ((int64_t)d.nz * fd.t);
    j = fd.high ? d.ny - fd.t + l : l;
  } else {
    l = p % fd.t; j = (p / fd.t) % d.ny; i = p / ((int64_t)fd.t * d.ny);
    k = fd.high ? d.nz - fd.t + l : l;
  }
  const int64_t cell = ((int64_t)i * d.ny + j) * d.nz + k;
  T kc[3], v[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    kc[c] = sc * mat[c * d.cs + cell];
    v[c] = F[c * d.cs + cell];
  }
  if (IS_E)
    face_update_E(fd, d, G, cell, l, p, kc, v);
  else
    face_update_H(fd, d, G, cell, l, p, kc, v);
  const int c0 = face_comp(fd.axis, 0), c1 = face_comp(fd.axis, 1);
  F[c0 * d.cs + cell] = v[c0];
  F[c1 * d.cs + cell] = v[c1];
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
  return a;
}

template <typename T>
void launch_main(const FdtdPlan* p, const MainArgs<T>& a, bool is_H, hipStream_t st) {
  const int bz = 1 << a.bz_shift, by = 256 >> a.bz_shift;
  dim3 grid((p->d.nz + bz - 1) / bz, (p->d.ny + by - 1) / by, p->d.nx);
  if (is_H)
    hipLaunchKernelGGL(k_main_H<T>, grid, dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(k_main_E<T>, grid, dim3(256), 0, st, a);
}

template <typename T>
int launch_half(FdtdPlan* p, bool is_H, int64_t q, hipStream_t st) {
  const MainArgs<T> a = main_args<T>(p, is_H);
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
      const unsigned nb = (unsigned)((fd.slab + 255) / 256);
      if (is_H)
        hipLaunchKernelGGL((k_face<T, false>), dim3(nb), dim3(256), 0, st, p->d, a.sc, a.F, a.G, a.mat, fd);
      else
        hipLaunchKernelGGL((k_face<T, true>), dim3(nb), dim3(256), 0, st, p->d, a.sc, a.F, a.G, a.mat, fd);
    }
  }
  bool any_src = false;
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

int fdtd_bench_main_kernel(FdtdPlan* p, int is_H, int64_t reps, void* stream, float* ms) {
  if (int rc = check_plan(p)) return rc;
  if (reps < 1 || !ms) return fail(-1, "bad benchmark arguments");
  if (int rc = select_device(p)) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  hipEvent_t e0, e1;
  FDTD_HIP(hipEventCreate(&e0));
  FDTD_HIP(hipEventCreate(&e1));
  FDTD_HIP(hipEventRecord(e0, st));
  for (int64_t r = 0; r < reps; ++r) {
    if (p->dtype == FDTD_F64)
      launch_main<double>(p, main_args<double>(p, is_H != 0), is_H != 0, st);
    else
      launch_main<float>(p, main_args<float>(p, is_H != 0), is_H != 0, st);
  }
  FDTD_HIP(hipGetLastError());
  FDTD_HIP(hipEventRecord(e1, st));
  FDTD_HIP(hipEventSynchronize(e1));
  FDTD_HIP(hipEventElapsedTime(ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 0;
}

}  // extern "C"
