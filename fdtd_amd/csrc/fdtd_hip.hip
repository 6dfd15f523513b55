// libfdtd_hip: host side of the C ABI declared in include/fdtd_hip.h (plan
// construction, launch sequencing, measurement hooks). The kernels are in
// fdtd_kernels.hpp.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "fdtd_hip.h"
#include "fdtd_kernels.hpp"

using namespace fdtd_dev;


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

// ------------------------------------------------------------------- host
struct FaceHost {
  int axis, high, t;
  std::vector<double> bE, cE, bH, cH;
  void* psiE;
  void* psiH;
};
struct ObjHost {
  int kind, x0, x1, y0, y1, z0, z1;
  int real;  // FDTD_OBJ_REAL_ARRAYS was given
  const void* inv;
  const void* absorb;
};
struct SrcHost {
  int kind, compE, compH;
  std::vector<int64_t> cells;
  std::vector<double> profile;
  double amplitude;
  // FDTD_SRC_SOFT
  double spacing = 0;
  int det = -1;
  void* vin = nullptr;
  void* vout = nullptr;
  int64_t cap = 0, count = 0;
  const int64_t* d_cells = nullptr;  // the cells in the plan's device table (set by finalize)
};
struct CurHost {
  std::vector<int32_t> pts;  // px[n], py[n], pz[n]
  double spacing;
  void* buf;
  int64_t cap, count;
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
  std::vector<CurHost> curs;
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
  CurDev* d_curs = nullptr;
  std::vector<FaceDev> h_faces;
  int n_fused = 0;           // leading faces folded into the main kernels
  std::vector<Bnd> stages;   // boundaries applied after the main kernel
  int64_t nE = 0, nH = 0;    // half steps taken (detector slot clocks)
  int cx = 1;                // x planes per workgroup of the main kernels
  int block_threads = 256;   // threads per workgroup of the main kernels
  bool fold_points = false;  // point sources added inside the E kernel
  bool nontemporal = true;   // read-once streams marked non-temporal
  bool cyclic = false;       // x-chunks dealt cyclically to the XCDs (large grids)
  bool flat = false;         // flat tiles: block_threads consecutive cells of the (y, z) plane
  // FDTD_F32C: 8-bit compensation terms of E and H
  void* E_lo = nullptr;
  void* H_lo = nullptr;
  // material palette (any dtype): one byte per cell + float64 (256, 3) tables
  const uint8_t* eps_id = nullptr;
  const uint8_t* mu_id = nullptr;
  const double* eps_tab = nullptr;
  const double* mu_tab = nullptr;
  // differentiation record of the next E half step (fdtd_run_recorded):
  // curl inside rec_db from the main kernel, E inside rec_pb from the tail
  void* rec_curl = nullptr;
  void* rec_plane = nullptr;
  Box rec_db{}, rec_pb{};
};

struct FdtdBatch {
  std::vector<FdtdPlan*> plans;
  int block_threads = 256, cx = 1, groups = 0;  // launch shape (groups 0: one per work item)
  void* dev = nullptr;        // sims_E, sims_H, first, q
  void* sims_E = nullptr;
  void* sims_H = nullptr;
  int64_t* first = nullptr;
  int32_t* item_sim = nullptr;
  int32_t* rot = nullptr;
  int64_t* q = nullptr;
  int64_t total = 0;
  bool cyclic = false;        // x-chunks dealt cyclically to the XCDs (BatchArgs)
  bool objs = false;          // some grid has objects: the E kernel variant with them
  // captured steps
  hipStream_t gstream = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipGraphExec_t gexec = nullptr;
  int64_t gsteps = 0;
};

struct FdtdAdjoint {
  FdtdPlan* plan = nullptr;
  const void* gcoef = nullptr;
  const void* dcoef = nullptr;
  void* lamE = nullptr;
  void* lamH = nullptr;
  void* psi = nullptr;
  Box db{}, pb{};
  double* grad = nullptr;
  // AdjSide table in device memory: [half step (H, E)][parity][axis][side]
  void* d_sides = nullptr;
  int parE = 0, parH = 0;
  int cx = 0;  // x planes per workgroup of the transposed kernels (0: the plan's)
  void* ybuf = nullptr;  // (3, N) Y buffer when the plan has anisotropic objects
};

namespace {

int check_plan(const FdtdPlan* p) {
  if (!p) return fail(-1, "null plan");
  if (!p->finalized) return fail(-2, "plan not finalized");
  return 0;
}

int cell_ok(const FdtdPlan* p, int64_t c) {
  return c >= 0 && c < p->d.cs && c % p->d.px < (int64_t)p->d.ny * p->d.nz;  // not in a plane's padding
}

// cells of the grid (without the planes' padding)
int64_t grid_cells(const FdtdPlan* p) { return (int64_t)p->d.nx * p->d.ny * p->d.nz; }

// cells of one plane normal to `axis`
int64_t plane_cells(const FdtdPlan* p, int axis) {
  return grid_cells(p) / (axis == 0 ? p->d.nx : (axis == 1 ? p->d.ny : p->d.nz));
}

bool is_complex(const FdtdPlan* p) { return p->dtype == FDTD_C128 || p->dtype == FDTD_C64; }

// log2 of the lanes of a z-row of nz cells, one cell per lane (at most a
// wavefront)
int lanes_shift(int nz) {
  int s = 0;
  while ((1 << s) < nz && s < 6) ++s;
  return s;
}

void free_tables(FdtdPlan* p) {
  // kernels already queued may still read the old tables
  if (p->dev) (void)hipDeviceSynchronize();
  if (p->dev) (void)hipFree(p->dev);
  p->dev = nullptr;
  p->d_faces = nullptr;
  p->d_objs = nullptr;
  p->d_srcs = nullptr;
  p->d_dets = nullptr;
  p->d_curs = nullptr;
}

int select_device(const FdtdPlan* p) {
  int cur = -1;
  FDTD_HIP(hipGetDevice(&cur));
  if (cur != p->device) FDTD_HIP(hipSetDevice(p->device));
  return 0;
}

// Field and material policies of a plan (see fdtd_kernels.hpp).
template <class FS>
FS make_field(const FdtdPlan* p, bool H);
template <>
PlainField<double, double> make_field(const FdtdPlan* p, bool H) {
  return {reinterpret_cast<double*>(H ? p->H : p->E), p->d.cs};
}
template <>
PlainField<float, float> make_field(const FdtdPlan* p, bool H) {
  return {reinterpret_cast<float*>(H ? p->H : p->E), p->d.cs};
}
template <>
CompField make_field(const FdtdPlan* p, bool H) {
  return {reinterpret_cast<float*>(H ? p->H : p->E), p->d.cs};
}
template <>
PlainField<cplx<double>, cplx<double>> make_field(const FdtdPlan* p, bool H) {
  return {reinterpret_cast<cplx<double>*>(H ? p->H : p->E), p->d.cs};
}
template <>
PlainField<cplx<float>, cplx<float>> make_field(const FdtdPlan* p, bool H) {
  return {reinterpret_cast<cplx<float>*>(H ? p->H : p->E), p->d.cs};
}
template <class MS>
MS make_mat(const FdtdPlan* p, bool mu);
template <>
MatArray<double, cplx<double>> make_mat(const FdtdPlan* p, bool mu) {
  return {reinterpret_cast<const double*>(mu ? p->inv_mu : p->inv_eps), p->d.cs};
}
template <>
MatArray<float, cplx<float>> make_mat(const FdtdPlan* p, bool mu) {
  return {reinterpret_cast<const float*>(mu ? p->inv_mu : p->inv_eps), p->d.cs};
}
template <>
MatArray<double, double> make_mat(const FdtdPlan* p, bool mu) {
  return {reinterpret_cast<const double*>(mu ? p->inv_mu : p->inv_eps), p->d.cs};
}
template <>
MatArray<float, float> make_mat(const FdtdPlan* p, bool mu) {
  return {reinterpret_cast<const float*>(mu ? p->inv_mu : p->inv_eps), p->d.cs};
}
template <>
MatPalette<double> make_mat(const FdtdPlan* p, bool mu) {
  return {mu ? p->mu_id : p->eps_id, mu ? p->mu_tab : p->eps_tab};
}
template <>
MatPalette<float> make_mat(const FdtdPlan* p, bool mu) {
  return {mu ? p->mu_id : p->eps_id, mu ? p->mu_tab : p->eps_tab};
}

template <class FS, class MS>
MainArgs<FS, MS> main_args(const FdtdPlan* p, bool is_H) {
  MainArgs<FS, MS> a{};
  a.d = p->d;
  a.sc = (typename FS::R)p->sc;
  a.F = make_field<FS>(p, is_H);
  a.G = make_field<FS>(p, !is_H);
  a.M = make_mat<MS>(p, is_H);
  a.faces = p->d_faces;
  a.nfaces = p->n_fused;
  a.objs = p->d_objs;
  a.nobjs = is_H ? 0 : (int)p->objs.size();
  const int bzs = lanes_shift(p->d.nz);
  a.bz_shift = bzs;
  a.cx = p->cx;
  a.by = std::max(1, p->block_threads >> bzs);
  a.cyclic = p->cyclic;
  a.flat = p->flat ? p->block_threads : 0;
  a.npts = 0;
  return a;
}

// threads per workgroup of main-kernel arguments a
template <class FS, class MS>
unsigned block_of(const MainArgs<FS, MS>& a) {
  return a.flat ? (unsigned)a.flat : (unsigned)((1 << a.bz_shift) * a.by);
}

template <class FS, class MS, typename O>
void launch_main(const FdtdPlan* p, const MainArgs<FS, MS>& a, bool is_H, hipStream_t st) {
  const unsigned tiles = plane_tiles(p->d, a.bz_shift, a.by, a.flat);
  const unsigned nchunks = (unsigned)((p->d.nx + a.cx - 1) / a.cx);
  // x-chunks dealt cyclically to the 8 XCDs (xcd_chunk_tile): whole rounds of 8
  const dim3 grid(tiles * (a.cyclic ? (nchunks + 7) / 8 * 8 : nchunks)), block(block_of(a));
  if (a.rec) {  // recording forward pass of a gradient (plain storage)
    if (a.nobjs)
      hipLaunchKernelGGL((k_main_E<FS, MS, O, true, true, true>), grid, block, 0, st, a);
    else
      hipLaunchKernelGGL((k_main_E<FS, MS, O, false, true, true>), grid, block, 0, st, a);
  } else if (p->nontemporal) {
    if (is_H)
      hipLaunchKernelGGL((k_main_H<FS, MS, true>), grid, block, 0, st, a);
    else if (a.nobjs)
      hipLaunchKernelGGL((k_main_E<FS, MS, O, true, true, false>), grid, block, 0, st, a);
    else
      hipLaunchKernelGGL((k_main_E<FS, MS, O, false, true, false>), grid, block, 0, st, a);
  } else {
    if (is_H)
      hipLaunchKernelGGL((k_main_H<FS, MS, false>), grid, block, 0, st, a);
    else if (a.nobjs)
      hipLaunchKernelGGL((k_main_E<FS, MS, O, true, false, false>), grid, block, 0, st, a);
    else
      hipLaunchKernelGGL((k_main_E<FS, MS, O, false, false, false>), grid, block, 0, st, a);
  }
}

template <class FS, class MS, typename O>
int launch_half(FdtdPlan* p, bool is_H, int64_t q, hipStream_t st) {
  MainArgs<FS, MS> a = main_args<FS, MS>(p, is_H);
  if (!is_H && p->fold_points) {
    a.npts = (int32_t)p->srcs.size();
    for (int s = 0; s < a.npts; ++s) a.pt_cell[s] = p->srcs[s].cells[0];
    a.wave = p->wave;
    a.wave_cap = p->wave_cap;
    a.q = q;
  }
  if (!is_H && p->rec_curl) {
    a.rec = reinterpret_cast<typename FS::R*>(p->rec_curl);
    a.rb = p->rec_db;
  }
  launch_main<FS, MS, O>(p, a, is_H, st);
  for (const Bnd& b : p->stages) {
    if (b.periodic) {
      const int axis = b.index;
      const int n_axis = axis == 0 ? p->d.nx : (axis == 1 ? p->d.ny : p->d.nz);
      if (n_axis <= 1) continue;  // plane 0 is plane -1: the copy is the identity
      const int64_t n_plane = plane_cells(p, axis);
      hipLaunchKernelGGL(k_periodic<FS>, dim3((unsigned)((n_plane + 255) / 256)), dim3(256), 0, st, p->d, a.F,
                         axis, is_H ? 0 : 1);
    } else {
      const FaceDev& fd = p->h_faces[b.index];
      const dim3 nb((unsigned)((fd.slab + 255) / 256)), nt(256);
      switch (fd.axis * 2 + (is_H ? 1 : 0)) {
        case 0: hipLaunchKernelGGL((k_face<FS, MS, true, 0>), nb, nt, 0, st, p->d, a.sc, a.F, a.G, a.M, fd); break;
        case 1: hipLaunchKernelGGL((k_face<FS, MS, false, 0>), nb, nt, 0, st, p->d, a.sc, a.F, a.G, a.M, fd); break;
        case 2: hipLaunchKernelGGL((k_face<FS, MS, true, 1>), nb, nt, 0, st, p->d, a.sc, a.F, a.G, a.M, fd); break;
        case 3: hipLaunchKernelGGL((k_face<FS, MS, false, 1>), nb, nt, 0, st, p->d, a.sc, a.F, a.G, a.M, fd); break;
        case 4: hipLaunchKernelGGL((k_face<FS, MS, true, 2>), nb, nt, 0, st, p->d, a.sc, a.F, a.G, a.M, fd); break;
        default: hipLaunchKernelGGL((k_face<FS, MS, false, 2>), nb, nt, 0, st, p->d, a.sc, a.F, a.G, a.M, fd); break;
      }
    }
  }
  bool any_src = false;
  if (!p->fold_points)
    for (const SrcHost& s : p->srcs) any_src |= (!is_H || s.kind == FDTD_SRC_PLANE);
  const bool rec_plane = !is_H && p->rec_plane;
  if (any_src || !p->dets.empty() || rec_plane || (is_H && !p->curs.empty())) {
    TailArgs<FS> t{};
    t.d = p->d;
    t.F = a.F;
    t.is_H = is_H ? 1 : 0;
    t.nsrc = p->fold_points ? 0 : (int)p->srcs.size();
    if (rec_plane) {
      t.rec = reinterpret_cast<typename FS::R*>(p->rec_plane);
      t.rb = p->rec_pb;
    }
    t.ndet = (int)p->dets.size();
    t.srcs = p->d_srcs;
    t.dets = p->d_dets;
    t.wave = p->wave;
    t.wave_cap = p->wave_cap;
    t.q = q;
    t.counter = is_H ? p->nH : p->nE;
    t.counter_H = p->nH;
    t.ncur = (int32_t)p->curs.size();
    t.curs = p->d_curs;
    hipLaunchKernelGGL(k_tail<FS>, dim3(1), dim3(256), 0, st, t);
  }
  FDTD_HIP(hipGetLastError());
  if (is_H) {
    ++p->nH;
    for (DetHost& dh : p->dets) ++dh.countH;
    for (CurHost& ch : p->curs) ++ch.count;
  } else {
    ++p->nE;
    for (DetHost& dh : p->dets) ++dh.countE;
    for (SrcHost& sh : p->srcs)
      if (sh.kind == FDTD_SRC_SOFT) ++sh.count;
  }
  return 0;
}

// Calls F with the plan's policy types: F.template operator()<FS, MS, O>().
template <class Fn>
int with_policies(const FdtdPlan* p, Fn&& fn) {
  const bool pal = p->eps_id != nullptr;
  // complex fields: real grid materials, complex object arrays, no palette
  if (p->dtype == FDTD_C128)
    return fn.template operator()<PlainField<cplx<double>, cplx<double>>, MatArray<double, cplx<double>>, cplx<double>>();
  if (p->dtype == FDTD_C64)
    return fn.template operator()<PlainField<cplx<float>, cplx<float>>, MatArray<float, cplx<float>>, cplx<float>>();
  if (p->dtype == FDTD_F64)
    return pal ? fn.template operator()<PlainField<double, double>, MatPalette<double>, double>()
               : fn.template operator()<PlainField<double, double>, MatArray<double, double>, double>();
  if (p->dtype == FDTD_F32)
    return pal ? fn.template operator()<PlainField<float, float>, MatPalette<float>, float>()
               : fn.template operator()<PlainField<float, float>, MatArray<float, float>, float>();
  return pal ? fn.template operator()<CompField, MatPalette<double>, double>()
             : fn.template operator()<CompField, MatArray<double, double>, double>();
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
  for (size_t i = 0; i < p->curs.size(); ++i)
    if (p->curs[i].count + (e_only ? 0 : nsteps) > p->curs[i].cap)
      return fail(-4, "current detector " + std::to_string(i) + " buffer too small");
  for (const SrcHost& sh : p->srcs) {
    if (sh.kind != FDTD_SRC_SOFT) continue;
    if (sh.count + (is_H_only ? 0 : nsteps) > sh.cap) return fail(-4, "soft source voltage buffer too small");
    // the reference reads its detector's last current for q > 0 (sources.py:613-616)
    if (!is_H_only && nsteps > 0 && q0 > 0 && p->curs[sh.det].count == 0)
      return fail(-5, "soft source at time index > 0 before its current detector has a sample");
  }
  return 0;
}

int half_dispatch(FdtdPlan* p, bool is_H, int64_t q, hipStream_t st) {
  return with_policies(p, [&]<class FS, class MS, typename O>() { return launch_half<FS, MS, O>(p, is_H, q, st); });
}

int adj_check_plan(const FdtdPlan* p) {
  if (int rc = check_plan(p)) return rc;
  int side[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  for (const FaceHost& f : p->faces)
    if (++side[f.axis][f.high ? 1 : 0] > 1) return fail(-1, "adjoint needs at most one PML face per side");
  if (p->dtype != FDTD_F64 && p->dtype != FDTD_F32) return fail(-1, "adjoint needs FDTD_F64 or FDTD_F32 storage");
  // periodic boundaries are transposed as plane copies after the main
  // kernels; a PML face registered after one would be a separate stage
  for (const Bnd& b : p->stages)
    if (!b.periodic) return fail(-1, "adjoint needs every PML face registered before the periodic boundaries");
  // an anisotropic update adds to E; an absorbing one also scales it, so the
  // two do not compose in either order when they overlap
  auto overlap = [](const ObjHost& u, const ObjHost& v) {
    return std::max(u.x0, v.x0) < std::min(u.x1, v.x1) && std::max(u.y0, v.y0) < std::min(u.y1, v.y1) &&
           std::max(u.z0, v.z0) < std::min(u.z1, v.z1);
  };
  for (const ObjHost& o : p->objs)
    if (o.kind == FDTD_OBJ_ANISOTROPIC)
      for (const ObjHost& q : p->objs)
        if (q.kind == FDTD_OBJ_ABSORBING && overlap(o, q))
          return fail(-1, "adjoint does not support an anisotropic object overlapping an absorbing one");
  for (const SrcHost& s : p->srcs)
    if (s.kind == FDTD_SRC_SOFT) return fail(-1, "adjoint does not support soft sources");
  return 0;
}

// The hard (plane) sources overwrite E and H at their cells after each half
// step (sources.py:476-486), so those values do not depend on the state
// before: the transposed half step starts by zeroing the adjoint field there.
template <typename T>
void adj_mask_hard_sources(const FdtdPlan* p, T* lam, bool is_H, hipStream_t st) {
  for (const SrcHost& s : p->srcs) {
    if (s.kind != FDTD_SRC_PLANE) continue;
    const int n = (int)s.cells.size();
    const int64_t base = (int64_t)(is_H ? s.compH : s.compE) * p->d.cs;
    hipLaunchKernelGGL(k_adj_zero<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, lam + base, s.d_cells, n);
  }
}

// Transposes of the plan's periodic copies of one half step (the stages
// after the main kernel; adj_check_plan admits periodic stages only), in
// reverse registration order.
template <typename T>
void adj_periodic(const FdtdPlan* p, T* lam, bool is_H, hipStream_t st) {
  for (auto it = p->stages.rbegin(); it != p->stages.rend(); ++it) {
    const int axis = it->index;
    const int n_axis = axis == 0 ? p->d.nx : (axis == 1 ? p->d.ny : p->d.nz);
    if (n_axis <= 1) continue;  // the forward copy is the identity
    const int64_t n_plane = plane_cells(p, axis);
    hipLaunchKernelGGL(k_adj_periodic<T>, dim3((unsigned)((n_plane + 255) / 256)), dim3(256), 0, st, lam, p->d, axis,
                       is_H ? 0 : 1);
  }
}

bool box_ok(const FdtdPlan* p, const int32_t* b) {
  return b && b[0] >= 0 && b[0] < b[1] && b[1] <= p->d.nx && b[2] >= 0 && b[2] < b[3] && b[3] <= p->d.ny &&
         b[4] >= 0 && b[4] < b[5] && b[5] <= p->d.nz;
}

template <typename T>
AdjArgs<T> adj_args(const FdtdAdjoint* a) {
  const FdtdPlan* p = a->plan;
  AdjArgs<T> r{};
  r.d = p->d;
  r.sc = (T)p->sc;
  r.lamE = reinterpret_cast<T*>(a->lamE);
  r.lamH = reinterpret_cast<T*>(a->lamH);
  r.inv_eps = reinterpret_cast<const T*>(p->inv_eps);
  r.inv_mu = reinterpret_cast<const T*>(p->inv_mu);
  r.gcoef = reinterpret_cast<const T*>(a->gcoef);
  r.dcoef = reinterpret_cast<const T*>(a->dcoef);
  r.objs = p->d_objs;
  r.nobjs = (int32_t)p->objs.size();
  r.ybuf = reinterpret_cast<T*>(a->ybuf);
  r.db = a->db;
  r.pb = a->pb;
  r.grad = a->grad;
  int bzs = 0;
  while ((1 << bzs) < p->d.nz && bzs < 6) ++bzs;
  r.bz_shift = bzs;
  int cx = a->cx > 0 ? a->cx : p->cx;
  if (const char* env = getenv("FDTD_HIP_ADJ_CX")) cx = atoi(env);
  r.cx = std::max(1, std::min(cx, p->d.nx));
  r.by = std::max(1, p->block_threads >> bzs);
  return r;
}

// The CPML faces of the transposed half steps (H: psiH, bH, cH; E: psiE,
// bE, cE), for both parities: buffer `parity` is read, the other written.
// Face f's adjoint buffers in psi_adj: psiE[0], psiE[1], psiH[0], psiH[1],
// (2, slab) values each.
template <typename T>
std::vector<AdjSide<T>> adj_side_table(const FdtdPlan* p, void* psi_adj) {
  std::vector<AdjSide<T>> tab(2 * 2 * 6);
  T* base = reinterpret_cast<T*>(psi_adj);
  int64_t off = 0;
  for (size_t f = 0; f < p->faces.size(); ++f) {
    const FaceDev& fd = p->h_faces[f];
    for (int half = 0; half < 2; ++half) {  // 0: H, 1: E
      T* buf0 = base + off + (half ? 0 : 4 * fd.slab);
      for (int par = 0; par < 2; ++par) {
        AdjSide<T>& e = tab[(half * 2 + par) * 6 + fd.axis * 2 + (fd.high ? 1 : 0)];
        e.pin = buf0 + 2 * par * fd.slab;
        e.pout = buf0 + 2 * (1 - par) * fd.slab;
        e.b = half ? fd.bE : fd.bH;
        e.c = half ? fd.cE : fd.cH;
        e.slab = fd.slab;
        e.t = fd.t;
      }
    }
    off += 8 * fd.slab;
  }
  return tab;
}

template <typename T>
const AdjSide<T>* adj_sides(const FdtdAdjoint* a, bool is_H, int parity) {
  return reinterpret_cast<const AdjSide<T>*>(a->d_sides) + ((is_H ? 0 : 2) + parity) * 6;
}

template <typename T>
int adjoint_steps(FdtdAdjoint* a, int64_t nsteps, const T* curl, const T* plane, hipStream_t st) {
  const FdtdPlan* p = a->plan;
  const int64_t nd = box_cells(a->db), np = box_cells(a->pb);
  AdjArgs<T> args = adj_args<T>(a);
  const int bz = 1 << args.bz_shift, by = args.by;
  const unsigned tiles = (unsigned)(((p->d.nz + bz - 1) / bz) * ((p->d.ny + by - 1) / by));
  const dim3 grid(tiles * (unsigned)((p->d.nx + args.cx - 1) / args.cx)), block(bz * by);
  const dim3 ygrid((unsigned)((p->d.cs + 255) / 256));
  // no thread past the (y, z) edges: k_adj_H without its store guard
  // (FDTD_HIP_ADJ_FULL=0 keeps the guard, to check that both forms agree)
  const char* full_env = getenv("FDTD_HIP_ADJ_FULL");
  const bool full = p->d.nz % bz == 0 && p->d.ny % by == 0 && !(full_env && atoi(full_env) == 0);
  for (int64_t s = nsteps - 1; s >= 0; --s) {
    // each half step transposed in reverse: hard sources, then the periodic
    // copies (reverse registration order), then the main update
    args.plane = plane + s * 3 * np;
    args.side = adj_sides<T>(a, true, a->parH);
    adj_mask_hard_sources<T>(p, args.lamH, true, st);
    adj_periodic<T>(p, args.lamH, true, st);
    if (full)
      hipLaunchKernelGGL((k_adj_H<T, true>), grid, block, 0, st, args);
    else
      hipLaunchKernelGGL((k_adj_H<T, false>), grid, block, 0, st, args);
    a->parH ^= 1;
    args.curl = curl + s * 3 * nd;
    args.side = adj_sides<T>(a, false, a->parE);
    adj_mask_hard_sources<T>(p, args.lamE, false, st);
    adj_periodic<T>(p, args.lamE, false, st);
    if (args.ybuf) {
      hipLaunchKernelGGL(k_adj_Y<T>, ygrid, dim3(256), 0, st, args);
      hipLaunchKernelGGL((k_adj_E<T, true>), grid, block, 0, st, args);
    } else {
      hipLaunchKernelGGL((k_adj_E<T, false>), grid, block, 0, st, args);
    }
    a->parE ^= 1;
  }
  FDTD_HIP(hipGetLastError());
  return 0;
}

// Forward steps with the records written by the main E kernel (curl) and
// the tail (E inside the objective box); curl may be null (objective only).
template <typename T>
int recorded_steps(FdtdAdjoint* a, int64_t q0, int64_t nsteps, T* curl, T* plane, hipStream_t st) {
  FdtdPlan* p = a->plan;
  const int64_t nd = box_cells(a->db), np = box_cells(a->pb);
  p->rec_db = a->db;
  p->rec_pb = a->pb;
  int rc = 0;
  for (int64_t s = 0; s < nsteps && rc == 0; ++s) {
    p->rec_curl = curl ? curl + s * 3 * nd : nullptr;
    p->rec_plane = plane + s * 3 * np;
    rc = half_dispatch(p, false, q0 + s, st);
    p->rec_curl = nullptr;
    p->rec_plane = nullptr;
    if (rc == 0) rc = half_dispatch(p, true, q0 + s, st);
  }
  return rc;
}

// ------------------------------------------------------------------ batch
int batch_check_plan(const FdtdPlan* p, const FdtdPlan* p0) {
  if (int rc = check_plan(p)) return rc;
  if (p->dtype != p0->dtype) return fail(-1, "batched grids must share one storage type");
  if (is_complex(p)) return fail(-1, "batched grids cannot have complex fields");
  if ((p->eps_id != nullptr) != (p0->eps_id != nullptr)) return fail(-1, "batched grids must all use a palette or none");
  if (!p->stages.empty()) return fail(-1, "batched grids cannot have periodic boundaries");
  if (!p->dets.empty()) return fail(-1, "batched grids cannot have detectors");
  if (!p->srcs.empty() && !p->fold_points) return fail(-1, "batched grids may have at most 4 point sources only");
  if (p->device != p0->device) return fail(-1, "batched grids must be on one device");
  return 0;
}

void batch_free_graph(FdtdBatch* b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
  b->gsteps = 0;
}

// Whether a batch deals its x-chunks cyclically to the XCDs: as for one grid,
// once the fields outgrow the caches (16.7M cells in all), and when its grids
// average 1M cells or more. Contiguous runs of large grids leave some XCDs
// working while others wait (the grids' CPML slabs and unequal shapes make
// tiles cost unequally); grids of a few MB keep their x-halo planes in the L2
// of the one XCD that a contiguous run gives them. FDTD_HIP_CYCLIC overrides.
bool batch_cyclic(const std::vector<FdtdPlan*>& plans) {
  int64_t cells = 0;
  for (const FdtdPlan* p : plans) cells += grid_cells(p);
  bool cyc = cells >= (int64_t)1 << 24 && cells / (int64_t)plans.size() >= (int64_t)1 << 20;
  if (const char* env = getenv("FDTD_HIP_CYCLIC")) cyc = atoi(env) != 0;
  return cyc;
}

// Per-grid kernel arguments of a batch for a launch shape: the E and H main
// arguments of every plan, first[s], the first position of grid s in the
// work list (first[n] = the number of positions), and rot[s], the XCD of grid
// s's first chunk under cyclic dealing (the chunks of all grids in turn dealt
// to XCD 0, 1, ..., 7, 0, ...; a grid spans whole rounds of 8 chunks).
template <class FS, class MS>
void batch_tables(const std::vector<FdtdPlan*>& plans, int block_threads, int cx, bool cyclic,
                  std::vector<MainArgs<FS, MS>>& sE, std::vector<MainArgs<FS, MS>>& sH, std::vector<int64_t>& first,
                  std::vector<int32_t>& rot, bool& objs) {
  using MA = MainArgs<FS, MS>;
  const size_t n = plans.size();
  sE.assign(n, MA{});
  sH.assign(n, MA{});
  first.assign(n + 1, 0);
  rot.assign(n, 0);
  objs = false;
  int64_t chunks = 0;  // chunks of the grids before this one
  for (size_t i = 0; i < n; ++i) {
    FdtdPlan* p = plans[i];
    MA aE = main_args<FS, MS>(p, false), aH = main_args<FS, MS>(p, true);
    const int bzs = lanes_shift(p->d.nz);
    for (MA* a : {&aE, &aH}) {
      a->cx = std::max(1, std::min(cx, p->d.nx));
      a->by = std::max(1, block_threads >> bzs);
      a->flat = p->flat ? block_threads : 0;
    }
    if (p->fold_points) {
      aE.npts = (int32_t)p->srcs.size();
      for (int s = 0; s < aE.npts; ++s) aE.pt_cell[s] = p->srcs[s].cells[0];
      aE.wave = p->wave;
      aE.wave_cap = p->wave_cap;
    }
    objs = objs || aE.nobjs > 0;
    sE[i] = aE;
    sH[i] = aH;
    const int64_t per = plane_tiles(p->d, aE.bz_shift, aE.by, aE.flat), nch = (p->d.nx + aE.cx - 1) / aE.cx;
    rot[i] = (int32_t)(chunks % 8);
    const int64_t span = cyclic ? 8 * per * ((rot[i] + nch + 7) / 8) : per * nch;
    chunks += nch;
    first[i + 1] = first[i] + span;
  }
}

// Host image of a batch's device tables for a launch shape: the E and H
// argument tables, first[], rot[], the time index slot and item_sim[], at the
// byte offsets below (each 256-byte aligned).
struct BatchImage {
  std::vector<unsigned char> bytes;
  size_t oE = 0, oH = 0, oF = 0, oR = 0, oQ = 0, oI = 0;
  int64_t total = 0;  // positions of the work list
  bool cyclic = false;
  bool objs = false;
};

int batch_image(const std::vector<FdtdPlan*>& plans, int block_threads, int cx, BatchImage& img) {
  const size_t n = plans.size();
  return with_policies(plans[0], [&]<class FS, class MS, typename O>() -> int {
    using MA = MainArgs<FS, MS>;
    std::vector<MA> sE, sH;
    std::vector<int64_t> first;
    std::vector<int32_t> rot;
    img.cyclic = batch_cyclic(plans);
    batch_tables<FS, MS>(plans, block_threads, cx, img.cyclic, sE, sH, first, rot, img.objs);
    img.total = first[n];
    if (img.total >= (int64_t)1 << 31) return fail(-1, "batch has too many work items");
    img.oE = 0;
    img.oH = align_up(n * sizeof(MA));
    img.oF = img.oH + align_up(n * sizeof(MA));
    img.oR = img.oF + align_up((n + 1) * sizeof(int64_t));
    img.oQ = img.oR + align_up(n * sizeof(int32_t));
    img.oI = img.oQ + 256;
    img.bytes.assign(img.oI + align_up(img.total * sizeof(int32_t)), 0);
    unsigned char* h = img.bytes.data();
    std::memcpy(h + img.oE, sE.data(), n * sizeof(MA));
    std::memcpy(h + img.oH, sH.data(), n * sizeof(MA));
    std::memcpy(h + img.oF, first.data(), (n + 1) * sizeof(int64_t));
    std::memcpy(h + img.oR, rot.data(), n * sizeof(int32_t));
    int32_t* item_sim = reinterpret_cast<int32_t*>(h + img.oI);
    for (size_t i = 0; i < n; ++i)
      for (int64_t t = first[i]; t < first[i + 1]; ++t) item_sim[t] = (int32_t)i;
    return 0;
  });
}

// (Re)build the per-grid argument tables for the current launch shape.
int batch_upload(FdtdBatch* b) {
  FdtdPlan* p0 = b->plans[0];
  if (int rc = select_device(p0)) return rc;
  batch_free_graph(b);
  if (b->dev) {
    (void)hipDeviceSynchronize();
    (void)hipFree(b->dev);
    b->dev = nullptr;
  }
  BatchImage img;
  if (int rc = batch_image(b->plans, b->block_threads, b->cx, img)) return rc;
  b->total = img.total;
  b->cyclic = img.cyclic;
  b->objs = img.objs;
  FDTD_HIP(hipMalloc(&b->dev, img.bytes.size()));
  FDTD_HIP(hipMemcpy(b->dev, img.bytes.data(), img.bytes.size(), hipMemcpyHostToDevice));
  char* base = reinterpret_cast<char*>(b->dev);
  b->sims_E = base + img.oE;
  b->sims_H = base + img.oH;
  b->first = reinterpret_cast<int64_t*>(base + img.oF);
  b->rot = reinterpret_cast<int32_t*>(base + img.oR);
  b->q = reinterpret_cast<int64_t*>(base + img.oQ);
  b->item_sim = reinterpret_cast<int32_t*>(base + img.oI);
  return 0;
}

int batch_groups(const FdtdBatch* b) {
  if (b->groups > 0 && b->groups < b->total) return b->groups;
  return (int)b->total;  // one workgroup per work item
}

// Device tables of a batched launch (see BatchArgs).
struct BatchTables {
  const void* sims_E;
  const void* sims_H;
  const int64_t* first;
  const int32_t* item_sim;
  const int32_t* rot;
  int64_t total;
  int64_t* q;
  bool cyclic;
};

// nsteps batched steps on stream st (the time index already set on the
// device), `groups` workgroups of `block_threads`
int batch_launch_tables(const FdtdPlan* p0, const BatchTables& t, bool objs, int groups, int block_threads,
                        int64_t nsteps, hipStream_t st) {
  const dim3 grid((unsigned)groups), block((unsigned)block_threads);
  return with_policies(p0, [&]<class FS, class MS, typename O>() -> int {
    BatchArgs<FS, MS> aE{reinterpret_cast<const MainArgs<FS, MS>*>(t.sims_E), t.first, t.item_sim, t.rot, t.total, t.q,
                         t.cyclic ? 1 : 0};
    BatchArgs<FS, MS> aH = aE;
    aH.sims = reinterpret_cast<const MainArgs<FS, MS>*>(t.sims_H);
    const bool loop = groups < t.total;  // a persistent grid of workgroups
    for (int64_t s = 0; s < nsteps; ++s) {
      if (objs && loop)
        hipLaunchKernelGGL((k_batch_E<FS, MS, O, true, true, true>), grid, block, 0, st, aE);
      else if (objs)
        hipLaunchKernelGGL((k_batch_E<FS, MS, O, true, true, false>), grid, block, 0, st, aE);
      else if (loop)
        hipLaunchKernelGGL((k_batch_E<FS, MS, O, false, true, true>), grid, block, 0, st, aE);
      else
        hipLaunchKernelGGL((k_batch_E<FS, MS, O, false, true, false>), grid, block, 0, st, aE);
      if (loop)
        hipLaunchKernelGGL((k_batch_H<FS, MS, true, true>), grid, block, 0, st, aH);
      else
        hipLaunchKernelGGL((k_batch_H<FS, MS, true, false>), grid, block, 0, st, aH);
    }
    FDTD_HIP(hipGetLastError());
    return 0;
  });
}

int batch_launch(FdtdBatch* b, int64_t nsteps, hipStream_t st) {
  const BatchTables t{b->sims_E, b->sims_H, b->first, b->item_sim, b->rot, b->total, b->q, b->cyclic};
  return batch_launch_tables(b->plans[0], t, b->objs, batch_groups(b), b->block_threads, nsteps, st);
}

// Pointers to the tables of an image placed at device address base.
BatchTables image_tables(const BatchImage& img, void* base) {
  char* c = reinterpret_cast<char*>(base);
  return BatchTables{c + img.oE,
                     c + img.oH,
                     reinterpret_cast<const int64_t*>(c + img.oF),
                     reinterpret_cast<const int32_t*>(c + img.oI),
                     reinterpret_cast<const int32_t*>(c + img.oR),
                     img.total,
                     reinterpret_cast<int64_t*>(c + img.oQ),
                     img.cyclic};
}

// ----------------------------------------------- batch without allocation
// fdtd_batch_run_static keeps a batch's tables in this array instead of in
// memory it allocates. Each GPU the library runs on holds its own copy of
// the array until the process ends. In host code `scratch` names no device
// address: the address of the current device's copy comes from
// hipGetSymbolAddress (static_scratch below), and only that is passed on.
constexpr size_t kStaticBatchBytes = FDTD_STATIC_BATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kStaticBatchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch array (device `dev` must be
// current); looked up once per device.
int static_scratch(int dev, void** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (dev < 0 || dev >= kMaxDevices) return fail(-1, "device index out of range for the static tables");
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    FDTD_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(-3, "hipGetSymbolAddress returned no address for the static tables");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

int batch_room(const std::vector<FdtdPlan*>& plans, int64_t q0, int64_t nsteps) {
  for (FdtdPlan* p : plans)
    if (int rc = check_step_room(p, q0, nsteps, false, false)) return rc;
  return 0;
}

void batch_advance(const std::vector<FdtdPlan*>& plans, int64_t nsteps) {
  for (FdtdPlan* p : plans) {
    p->nE += nsteps;
    p->nH += nsteps;
  }
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
  if (dtype != FDTD_F64 && dtype != FDTD_F32 && dtype != FDTD_F32C && dtype != FDTD_C128 && dtype != FDTD_C64)
    return fail(-1, "unknown dtype");
  if (nx < 1 || ny < 1 || nz < 1) return fail(-1, "grid dimensions must be positive");
  if (nx > 65535) return fail(-1, "nx above 65535 is not supported");
  if (!E || !H || !inv_eps || !inv_mu) return fail(-1, "null field or material pointer");
  FdtdPlan* p = new FdtdPlan();
  p->dtype = dtype;
  p->d.nx = nx;
  p->d.ny = ny;
  p->d.nz = nz;
  p->d.px = (int64_t)ny * nz;
  p->d.cs = (int64_t)nx * p->d.px;
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

int fdtd_plan_set_pitch(FdtdPlan* p, int64_t px) {
  if (!p) return fail(-1, "null plan");
  if (px < (int64_t)p->d.ny * p->d.nz) return fail(-1, "the x-plane pitch must be at least ny*nz");
  if (!p->faces.empty() || !p->objs.empty() || !p->srcs.empty() || !p->dets.empty() || !p->curs.empty() ||
      !p->order.empty() || p->E_lo || p->eps_id)
    return fail(-1, "set the pitch before anything is added to the plan");
  p->d.px = px;
  p->d.cs = (int64_t)p->d.nx * px;
  p->finalized = false;
  return 0;
}

int fdtd_plan_set_compensation(FdtdPlan* p, void* E_lo, void* H_lo) {
  if (!p) return fail(-1, "null plan");
  if (p->dtype != FDTD_F32C) return fail(-1, "compensation arrays belong to FDTD_F32C plans");
  if (!E_lo || !H_lo) return fail(-1, "null compensation array");
  if (E_lo != static_cast<char*>(p->E) + 3 * sizeof(float) || H_lo != static_cast<char*>(p->H) + 3 * sizeof(float))
    return fail(-1, "FDTD_F32C compensation words must be the 4th float of the E and H records");
  p->E_lo = E_lo;
  p->H_lo = H_lo;
  p->finalized = false;
  return 0;
}

int fdtd_plan_set_material_palette(FdtdPlan* p, const uint8_t* eps_id, const double* eps_table,
                                   const uint8_t* mu_id, const double* mu_table) {
  if (!p) return fail(-1, "null plan");
  const bool any = eps_id || eps_table || mu_id || mu_table;
  if (any && !(eps_id && eps_table && mu_id && mu_table)) return fail(-1, "palette needs ids and tables for both");
  if (any && is_complex(p)) return fail(-1, "no material palette on the complex-field path");
  p->eps_id = eps_id;
  p->eps_tab = eps_table;
  p->mu_id = mu_id;
  p->mu_tab = mu_table;
  p->finalized = false;
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
  const int real = (kind & FDTD_OBJ_REAL_ARRAYS) ? 1 : 0;
  kind &= ~FDTD_OBJ_REAL_ARRAYS;
  if (kind < FDTD_OBJ_PLAIN || kind > FDTD_OBJ_ANISOTROPIC) return fail(-1, "unknown object kind");
  if (x0 < 0 || y0 < 0 || z0 < 0 || x1 > p->d.nx || y1 > p->d.ny || z1 > p->d.nz || x0 >= x1 || y0 >= y1 ||
      z0 >= z1)
    return fail(-1, "object box outside the grid or empty");
  if (!inv) return fail(-1, "null object permittivity");
  if (kind == FDTD_OBJ_ABSORBING && !absorption) return fail(-1, "null absorption factor");
  p->objs.push_back({kind, x0, x1, y0, y1, z0, z1, real, inv, absorption});
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

int fdtd_plan_add_current_detector(FdtdPlan* p, int n, const int32_t* px, const int32_t* py, const int32_t* pz,
                                   double spacing, void* bufI, int64_t capacity, int64_t count) {
  if (!p) return fail(-1, "null plan");
  if (n < 1 || !px || !py || !pz || !bufI) return fail(-1, "current detector without points or buffer");
  if (is_complex(p)) return fail(-1, "current detectors are not supported on the complex-field path");
  if (count < 0 || count > capacity) return fail(-1, "bad current detector count");
  CurHost c;
  c.pts.resize(3 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    // numpy indexing: -n <= index < n, and the loop reads index - 1
    if (px[i] < 0 || px[i] >= p->d.nx || py[i] < 0 || py[i] >= p->d.ny || pz[i] < 0 || pz[i] >= p->d.nz)
      return fail(-1, "current detector point outside the grid");
    c.pts[i] = px[i];
    c.pts[n + i] = py[i];
    c.pts[2 * n + i] = pz[i];
  }
  c.spacing = spacing;
  c.buf = bufI;
  c.cap = capacity;
  c.count = count;
  p->curs.push_back(std::move(c));
  p->finalized = false;
  return 0;
}

int fdtd_plan_add_soft_source(FdtdPlan* p, int64_t cell, double impedance, double spacing, int detector, void* vin,
                              void* vout, int64_t capacity, int64_t count) {
  if (!p) return fail(-1, "null plan");
  if (!cell_ok(p, cell)) return fail(-1, "source cell outside the grid");
  if (!vin || !vout || count < 0 || count > capacity) return fail(-1, "bad soft source buffers");
  if (detector < 0) return fail(-1, "bad current detector index");
  if (is_complex(p)) return fail(-1, "soft sources are not supported on the complex-field path");
  SrcHost s;
  s.kind = FDTD_SRC_SOFT;
  s.compE = 2;
  s.compH = 0;
  s.cells.assign(1, cell);
  s.amplitude = impedance;
  s.spacing = spacing;
  s.det = detector;
  s.vin = vin;
  s.vout = vout;
  s.cap = capacity;
  s.count = count;
  p->srcs.push_back(std::move(s));
  p->finalized = false;
  return 0;
}

int fdtd_plan_current_detector_count(const FdtdPlan* p, int index, int64_t* count) {
  if (!p || !count) return fail(-1, "null argument");
  if (index < 0 || index >= (int)p->curs.size()) return fail(-1, "current detector index out of range");
  *count = p->curs[index].count;
  return 0;
}

int fdtd_plan_soft_source_count(const FdtdPlan* p, int index, int64_t* count) {
  if (!p || !count) return fail(-1, "null argument");
  int k = 0;
  for (const SrcHost& sh : p->srcs)
    if (sh.kind == FDTD_SRC_SOFT && k++ == index) {
      *count = sh.count;
      return 0;
    }
  return fail(-1, "soft source index out of range");
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
  if (p->dtype == FDTD_F32C && (!p->E_lo || !p->H_lo)) return fail(-1, "FDTD_F32C plan without compensation arrays");
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
    const int bzs = lanes_shift(p->d.nz);
    const int by = std::max(1, p->block_threads >> bzs);
    // flat tiles (a workgroup's cells consecutive in the (y, z) plane): no
    // lanes idle on uneven z edges, and a workgroup streams one contiguous
    // run of each array; measured on MI355X at 256^3 they beat tiles of four
    // 64-cell z-rows by 4-9 % in every storage. Planes of 2^31 cells and more
    // keep row tiles.
    const int64_t plane = (int64_t)p->d.ny * p->d.nz;
    p->flat = true;
    if (const char* env = getenv("FDTD_HIP_FLAT")) p->flat = atoi(env) != 0;
    p->flat = p->flat && plane < ((int64_t)1 << 31);
    const int64_t tiles = plane_tiles(p->d, bzs, by, p->flat ? p->block_threads : 0);
    // x planes per workgroup, measured on MI355X with flat tiles at 256^3 and
    // 512x512x384: float64 is fastest at 1 (3-4 % ahead of 2), float32 and the
    // compensated storage at 2 (5-6 % ahead of 1 and 3); every storage at 1 on
    // grids of ~1M cells
    const int64_t work = tiles * p->d.nx;
    int64_t cx = p->dtype == FDTD_F64 ? 1 : std::min<int64_t>(2, work / 32768);
    if (const char* env = getenv("FDTD_HIP_CX")) cx = atoi(env);
    if (const char* env = getenv("FDTD_HIP_NT")) p->nontemporal = atoi(env) != 0;
    // cyclic dealing pays once the fields outgrow the caches (16.7M cells = 256^3)
    p->cyclic = grid_cells(p) >= (int64_t)1 << 24;
    if (const char* env = getenv("FDTD_HIP_CYCLIC")) p->cyclic = atoi(env) != 0;
    p->cx = (int)std::max<int64_t>(1, std::min<int64_t>(cx, p->d.nx));
  }
  // point sources fold into the E kernel when nothing runs between the main
  // kernel and the sources (no periodic or staged face) and no detector samples
  p->fold_points = !p->srcs.empty() && (int)p->srcs.size() <= FOLDED_POINTS && p->dets.empty() && p->curs.empty();
  for (const SrcHost& s : p->srcs) p->fold_points = p->fold_points && s.kind == FDTD_SRC_POINT;
  for (const Bnd& b : p->order) p->fold_points = p->fold_points && !b.periodic;
  for (const SrcHost& sh : p->srcs)
    if (sh.kind == FDTD_SRC_SOFT && (sh.det < 0 || sh.det >= (int)p->curs.size()))
      return fail(-1, "soft source refers to a missing current detector");
  // byte layout of the table block
  const size_t nf = p->faces.size(), no = p->objs.size(), ns = p->srcs.size(), nd = p->dets.size(),
               nc = p->curs.size();
  size_t off = 0;
  const size_t o_faces = off;
  off = align_up(off + nf * sizeof(FaceDev));
  const size_t o_objs = off;
  off = align_up(off + no * sizeof(ObjDev));
  const size_t o_srcs = off;
  off = align_up(off + ns * sizeof(SrcDev));
  const size_t o_dets = off;
  off = align_up(off + nd * sizeof(DetDev));
  const size_t o_curs = off;
  off = align_up(off + nc * sizeof(CurDev));
  std::vector<size_t> o_cpts(nc);
  for (size_t i = 0; i < nc; ++i) {
    o_cpts[i] = off;
    off = align_up(off + p->curs[i].pts.size() * sizeof(int32_t));
  }
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
    fd.slab = plane_cells(p, f.axis) * f.t;
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
    ObjDev od{o.kind, o.x0, o.x1, o.y0, o.y1, o.z0, o.z1, o.real, o.inv, o.absorb};
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
    p->srcs[i].d_cells = sd.cells;
    sd.profile = s.profile.empty() ? nullptr : reinterpret_cast<const double*>(base + o_sprof[i]);
    sd.amplitude = s.amplitude;
    sd.spacing = s.spacing;
    sd.det = s.det;
    sd.vin = s.vin;
    sd.vout = s.vout;
    sd.slot0 = p->nE - s.count;
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
  for (size_t i = 0; i < nc; ++i) {
    const CurHost& ch = p->curs[i];
    std::memcpy(host.data() + o_cpts[i], ch.pts.data(), ch.pts.size() * sizeof(int32_t));
    CurDev cd{};
    cd.n = (int32_t)(ch.pts.size() / 3);
    cd.pts = reinterpret_cast<const int32_t*>(base + o_cpts[i]);
    cd.spacing = ch.spacing;
    cd.buf = ch.buf;
    cd.slot0 = p->nH - ch.count;
    std::memcpy(host.data() + o_curs + i * sizeof(CurDev), &cd, sizeof(CurDev));
  }
  FDTD_HIP(hipMemcpy(p->dev, host.data(), total, hipMemcpyHostToDevice));
  p->d_faces = reinterpret_cast<FaceDev*>(base + o_faces);
  p->d_objs = reinterpret_cast<ObjDev*>(base + o_objs);
  p->d_srcs = reinterpret_cast<SrcDev*>(base + o_srcs);
  p->d_dets = reinterpret_cast<DetDev*>(base + o_dets);
  p->d_curs = reinterpret_cast<CurDev*>(base + o_curs);
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

int fdtd_check_finite(const FdtdPlan* p, int32_t* flag, void* stream) {
  if (!p) return fail(-1, "null plan");
  if (!flag) return fail(-1, "null flag");
  if (int rc = select_device(p)) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const int64_t cs = p->d.cs;
  const dim3 grid(1024), block(256);
  for (void* f : {p->E, p->H}) {
    switch (p->dtype) {
      case FDTD_F64:
        hipLaunchKernelGGL(k_nonfinite<double>, grid, block, 0, st, (const double*)f, 3 * cs, 0, flag);
        break;
      case FDTD_C128:
        hipLaunchKernelGGL(k_nonfinite<double>, grid, block, 0, st, (const double*)f, 6 * cs, 0, flag);
        break;
      case FDTD_F32:
        hipLaunchKernelGGL(k_nonfinite<float>, grid, block, 0, st, (const float*)f, 3 * cs, 0, flag);
        break;
      case FDTD_C64:
        hipLaunchKernelGGL(k_nonfinite<float>, grid, block, 0, st, (const float*)f, 6 * cs, 0, flag);
        break;
      default:  // FDTD_F32C: (N, 4) records, the 4th float is the compensation word
        hipLaunchKernelGGL(k_nonfinite<float>, grid, block, 0, st, (const float*)f, 4 * cs, 4, flag);
        break;
    }
  }
  FDTD_HIP(hipGetLastError());
  return 0;
}

int fdtd_adjoint_psi_size(const FdtdPlan* p, int64_t* elements) {
  if (!p || !elements) return fail(-1, "null argument");
  int64_t n = 0;
  for (const FaceHost& f : p->faces)
    n += 8 * (plane_cells(p, f.axis) * f.t);
  *elements = n;
  return 0;
}

int fdtd_adjoint_create(FdtdPlan* p, const void* gcoef, const void* dcoef, void* lamE, void* lamH, void* psi_adj,
                        const int32_t design_box[6], const int32_t objective_box[6], double* grad,
                        FdtdAdjoint** out) {
  if (!out) return fail(-1, "null output pointer");
  *out = nullptr;
  if (int rc = adj_check_plan(p)) return rc;
  if (!gcoef || !lamE || !lamH || !grad) return fail(-1, "null adjoint array");
  if (!box_ok(p, design_box)) return fail(-1, "design box outside the grid or empty");
  if (!box_ok(p, objective_box)) return fail(-1, "objective box outside the grid or empty");
  if (!p->faces.empty() && !psi_adj) return fail(-1, "null adjoint CPML buffer");
  if (int rc = select_device(p)) return rc;
  FdtdAdjoint* a = new FdtdAdjoint();
  {
    std::vector<unsigned char> tab;
    if (p->dtype == FDTD_F64) {
      const auto t = adj_side_table<double>(p, psi_adj);
      tab.assign(reinterpret_cast<const unsigned char*>(t.data()),
                 reinterpret_cast<const unsigned char*>(t.data() + t.size()));
    } else {
      const auto t = adj_side_table<float>(p, psi_adj);
      tab.assign(reinterpret_cast<const unsigned char*>(t.data()),
                 reinterpret_cast<const unsigned char*>(t.data() + t.size()));
    }
    hipError_t e = hipMalloc(&a->d_sides, tab.size());
    if (e == hipSuccess) e = hipMemcpy(a->d_sides, tab.data(), tab.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      if (a->d_sides) (void)hipFree(a->d_sides);
      delete a;
      return fail(-3, std::string("adjoint face table: ") + hipGetErrorString(e));
    }
  }
  for (const ObjHost& o : p->objs) {
    if (o.kind != FDTD_OBJ_ANISOTROPIC) continue;
    const size_t bytes = (size_t)3 * p->d.cs * (p->dtype == FDTD_F64 ? 8 : 4);
    if (hipMalloc(&a->ybuf, bytes) != hipSuccess) {
      a->ybuf = nullptr;
      (void)hipFree(a->d_sides);
      delete a;
      return fail(-3, "adjoint Y buffer: out of device memory");
    }
    break;
  }
  a->plan = p;
  a->gcoef = gcoef;
  a->dcoef = dcoef;
  a->lamE = lamE;
  a->lamH = lamH;
  a->psi = psi_adj;
  a->db = Box{design_box[0], design_box[1], design_box[2], design_box[3], design_box[4], design_box[5]};
  a->pb = Box{objective_box[0], objective_box[1], objective_box[2], objective_box[3], objective_box[4],
              objective_box[5]};
  a->grad = grad;
  *out = a;
  return 0;
}

int fdtd_adjoint_destroy(FdtdAdjoint* a) {
  if (!a) return 0;
  if (a->d_sides || a->ybuf) (void)hipDeviceSynchronize();
  if (a->d_sides) (void)hipFree(a->d_sides);
  if (a->ybuf) (void)hipFree(a->ybuf);
  delete a;
  return 0;
}

int fdtd_run_recorded(FdtdAdjoint* a, int64_t q0, int64_t nsteps, void* curl_records, void* plane_records,
                      void* stream) {
  if (!a) return fail(-1, "null adjoint");
  FdtdPlan* p = a->plan;
  if (int rc = adj_check_plan(p)) return rc;
  if (nsteps < 0) return fail(-1, "negative step count");
  if (nsteps > 0 && !plane_records) return fail(-1, "null record buffer");
  if (int rc = check_step_room(p, q0, nsteps, false, false)) return rc;
  if (int rc = select_device(p)) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (p->dtype == FDTD_F64)
    return recorded_steps<double>(a, q0, nsteps, reinterpret_cast<double*>(curl_records),
                                  reinterpret_cast<double*>(plane_records), st);
  return recorded_steps<float>(a, q0, nsteps, reinterpret_cast<float*>(curl_records),
                               reinterpret_cast<float*>(plane_records), st);
}

int fdtd_adjoint_set_launch(FdtdAdjoint* a, int cx) {
  if (!a) return fail(-1, "null adjoint");
  if (cx < 0) return fail(-1, "cx must be 0 (the plan's) or positive");
  a->cx = cx;
  return 0;
}

int fdtd_adjoint_run(FdtdAdjoint* a, int64_t nsteps, const void* curl_records, const void* plane_records,
                     void* stream) {
  if (!a) return fail(-1, "null adjoint");
  if (int rc = adj_check_plan(a->plan)) return rc;
  if (nsteps < 0) return fail(-1, "negative step count");
  if (nsteps > 0 && (!curl_records || !plane_records)) return fail(-1, "null record buffer");
  if (int rc = select_device(a->plan)) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (a->plan->dtype == FDTD_F64)
    return adjoint_steps<double>(a, nsteps, reinterpret_cast<const double*>(curl_records),
                                 reinterpret_cast<const double*>(plane_records), st);
  return adjoint_steps<float>(a, nsteps, reinterpret_cast<const float*>(curl_records),
                              reinterpret_cast<const float*>(plane_records), st);
}

int fdtd_batch_create(int n, FdtdPlan* const* plans, FdtdBatch** out) {
  if (!out) return fail(-1, "null output pointer");
  *out = nullptr;
  if (n < 1 || !plans) return fail(-1, "a batch needs at least one plan");
  for (int i = 0; i < n; ++i)
    if (int rc = batch_check_plan(plans[i], plans[0])) return rc;
  FdtdBatch* b = new FdtdBatch();
  b->plans.assign(plans, plans + n);
  if (int rc = batch_upload(b)) {
    delete b;
    return rc;
  }
  *out = b;
  return 0;
}

int fdtd_batch_destroy(FdtdBatch* b) {
  if (!b) return 0;
  (void)hipDeviceSynchronize();
  batch_free_graph(b);
  if (b->gstream) (void)hipStreamDestroy(b->gstream);
  if (b->ev_in) (void)hipEventDestroy(b->ev_in);
  if (b->ev_out) (void)hipEventDestroy(b->ev_out);
  if (b->dev) (void)hipFree(b->dev);
  delete b;
  return 0;
}

int fdtd_batch_set_launch(FdtdBatch* b, int block_threads, int cx, int workgroups) {
  if (!b) return fail(-1, "null batch");
  if (block_threads != 256 && block_threads != 512 && block_threads != 1024)
    return fail(-1, "block_threads must be 256, 512 or 1024");
  if (cx < 1 || workgroups < 0) return fail(-1, "bad launch shape");
  b->block_threads = block_threads;
  b->cx = cx;
  b->groups = workgroups;
  return batch_upload(b);
}

int fdtd_batch_work_items(const FdtdBatch* b, int64_t* items) {
  if (!b || !items) return fail(-1, "null argument");
  *items = b->total;
  return 0;
}

int fdtd_batch_run(FdtdBatch* b, int64_t q0, int64_t nsteps, void* stream) {
  if (!b) return fail(-1, "null batch");
  if (nsteps < 0) return fail(-1, "negative step count");
  if (int rc = batch_room(b->plans, q0, nsteps)) return rc;
  if (int rc = select_device(b->plans[0])) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_set_time, dim3(1), dim3(64), 0, st, b->q, q0);
  if (int rc = batch_launch(b, nsteps, st)) return rc;
  batch_advance(b->plans, nsteps);
  return 0;
}

int fdtd_batch_run_graph(FdtdBatch* b, int64_t q0, int64_t nsteps, int64_t steps_per_graph, void* stream) {
  if (!b) return fail(-1, "null batch");
  if (nsteps < 0 || steps_per_graph < 1) return fail(-1, "bad step counts");
  if (int rc = batch_room(b->plans, q0, nsteps)) return rc;
  if (int rc = select_device(b->plans[0])) return rc;
  hipStream_t caller = reinterpret_cast<hipStream_t>(stream);
  if (!b->gstream) {
    FDTD_HIP(hipStreamCreateWithFlags(&b->gstream, hipStreamNonBlocking));
    FDTD_HIP(hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming));
    FDTD_HIP(hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming));
  }
  if (!b->gexec || b->gsteps != steps_per_graph) {
    batch_free_graph(b);
    hipGraph_t g = nullptr;
    FDTD_HIP(hipStreamBeginCapture(b->gstream, hipStreamCaptureModeThreadLocal));
    const int rc = batch_launch(b, steps_per_graph, b->gstream);
    const hipError_t e = hipStreamEndCapture(b->gstream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(-3, std::string("graph capture: ") + hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      b->gexec = nullptr;
      return fail(-3, std::string("graph instantiate: ") + hipGetErrorString(ei));
    }
    b->gsteps = steps_per_graph;
  }
  // the caller's stream -> the batch stream -> back
  hipLaunchKernelGGL(k_set_time, dim3(1), dim3(64), 0, caller, b->q, q0);
  FDTD_HIP(hipEventRecord(b->ev_in, caller));
  FDTD_HIP(hipStreamWaitEvent(b->gstream, b->ev_in, 0));
  const int64_t reps = nsteps / steps_per_graph, rest = nsteps % steps_per_graph;
  for (int64_t r = 0; r < reps; ++r) FDTD_HIP(hipGraphLaunch(b->gexec, b->gstream));
  if (rest)
    if (int rc = batch_launch(b, rest, b->gstream)) return rc;
  FDTD_HIP(hipEventRecord(b->ev_out, b->gstream));
  FDTD_HIP(hipStreamWaitEvent(caller, b->ev_out, 0));
  batch_advance(b->plans, nsteps);
  return 0;
}

int fdtd_batch_run_static(int n, FdtdPlan* const* plans, int block_threads, int cx, int workgroups, int64_t q0,
                          int64_t nsteps, void* stream) {
  if (n < 1 || !plans) return fail(-1, "a batch needs at least one plan");
  if (block_threads != 256 && block_threads != 512 && block_threads != 1024)
    return fail(-1, "block_threads must be 256, 512 or 1024");
  if (cx < 1 || workgroups < 0) return fail(-1, "bad launch shape");
  if (nsteps < 0) return fail(-1, "negative step count");
  for (int i = 0; i < n; ++i)
    if (int rc = batch_check_plan(plans[i], plans[0])) return rc;
  const std::vector<FdtdPlan*> pl(plans, plans + n);
  if (int rc = batch_room(pl, q0, nsteps)) return rc;
  if (int rc = select_device(pl[0])) return rc;
  BatchImage img;
  if (int rc = batch_image(pl, block_threads, cx, img)) return rc;
  if (img.bytes.size() > kStaticBatchBytes)
    return fail(-6, "the batch's tables need " + std::to_string(img.bytes.size()) + " bytes, more than the " +
                        std::to_string(kStaticBatchBytes) +
                        " of the static tables (FDTD_STATIC_BATCH_BYTES); use fdtd_batch_create for this batch");
  void* base = nullptr;
  if (int rc = static_scratch(pl[0]->device, &base)) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  // Stream-ordered after the work already queued on `stream` (an earlier
  // call's kernels may still read the tables); the wait also keeps the host
  // image alive until the copy has read it.
  FDTD_HIP(hipMemcpyAsync(base, img.bytes.data(), img.bytes.size(), hipMemcpyHostToDevice, st));
  FDTD_HIP(hipStreamSynchronize(st));
  const BatchTables t = image_tables(img, base);
  hipLaunchKernelGGL(k_set_time, dim3(1), dim3(64), 0, st, t.q, q0);
  const int groups = (workgroups > 0 && workgroups < img.total) ? workgroups : (int)img.total;
  if (int rc = batch_launch_tables(pl[0], t, img.objs, groups, block_threads, nsteps, st)) return rc;
  batch_advance(pl, nsteps);
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
      with_policies(p, [&]<class FS, class MS, typename O>() {
        launch_main<FS, MS, O>(p, main_args<FS, MS>(p, h != 0), h != 0, st);
        return 0;
      });
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
