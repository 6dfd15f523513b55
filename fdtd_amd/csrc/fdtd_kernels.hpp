// Device code of libfdtd_hip: fused Yee half-step kernels for MI355X (gfx950).
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
// The kernels are written once over two policies:
//   field storage  PlainField<S, R>  values stored as S, computed in R
//                                    (S = R = cplx<double/float> for the
//                                    complex-field path)
//                  CompField         float plus an 8-bit compensation term
//                                    (the next 8 mantissa bits: a 40-bit
//                                    format), computed in double
//   material       MatArray<S, R>    per-cell (3, N) array
//                  MatPalette<R>     one byte per cell indexing <= 256 float64
//                                    triples
//
// Kernels
//   k_main_E / k_main_H  curl, material update, objects and every CPML face
//                        registered before the first periodic boundary; a
//                        workgroup marches along x over a (y,z) tile.
//                        Memory-bound (~0.16 flop/B).
//   k_face               one CPML face over its slab, for faces registered
//                        after a periodic boundary (order matters there).
//   k_periodic           plane copy of a periodic boundary.
//   k_tail               one workgroup: sources (in order) then detectors.
//   k_batch_E / k_batch_H  the main kernels' tile bodies over the tiles of
//                        many grids (one launch per half step); k_set_time
//                        sets their device time index.
//   k_adj_H / k_adj_E    transposes of the H and E half steps (reverse mode),
//                        objective and design gradient fused.
//   k_copy               plain streaming copy (bandwidth reference).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fdtd_hip.h"

namespace fdtd_dev {

// ---------------------------------------------------------------- plan data
struct FaceDev {
  int32_t axis, high, t, pad;
  const double* bE;
  const double* cE;
  const double* bH;
  const double* cH;
  void* psiE;  // psi in the arithmetic type R of the plan
  void* psiH;
  int64_t slab;  // cells in the face slab
};

struct ObjDev {
  int32_t kind, x0, x1, y0, y1, z0, z1;
  int32_t real;        // the reference's arrays of this object are real (see absorb_ratio)
  const void* inv;     // in the plan's object type
  const void* absorb;
};

struct SrcDev {
  int32_t kind, n, compE, compH;
  const int64_t* cells;
  const double* profile;
  double amplitude;  // FDTD_SRC_SOFT: the impedance
  // FDTD_SRC_SOFT only
  double spacing;
  int32_t det, pad;  // current detector read
  void* vin;
  void* vout;
  int64_t slot0;     // slot written at E half-step counter c is c - slot0
};

// Current detector: n points (px, py, pz), loop currents after each H half step.
struct CurDev {
  int32_t n, pad;
  const int32_t* pts;  // px[n], py[n], pz[n]
  double spacing;
  void* buf;           // (capacity, n)
  int64_t slot0;       // slot written at H half-step counter c is c - slot0
};

struct DetDev {
  int32_t n, pad;
  const int64_t* cells;
  void* bufE;
  void* bufH;
  int64_t slot0E, slot0H;  // slot written at half-step counter c is c - slot0
};

// Cell (i, j, k) of a component is at i*px + j*nz + k: z fastest, x planes
// `px` cells apart (px >= ny*nz; the host may round a plane up so that every
// x plane starts on a cache line; the cells past ny*nz are then never read or
// written), components cs = nx*px apart.
struct Dims {
  int32_t nx, ny, nz, pad;
  int64_t cs;  // component stride nx*px
  int64_t px;  // x-plane pitch
};

__host__ __device__ __forceinline__ int64_t cell_of(const Dims& d, int64_t i, int64_t j, int64_t k) {
  return i * d.px + j * d.nz + k;
}

// ------------------------------------------------------- load/store helpers
// Loads and stores of the read-once streams (the updated field and the
// material); NT marks them non-temporal so the L2 keeps the other field,
// whose planes and rows are re-read as halos.
// Every array lives in global memory. Pointers read from tables in device
// memory (the per-grid arguments of the batched kernels, the face, object,
// source and detector tables) carry no address space, and accesses through
// them would be generic (flat) ones; gp() states the global space, which
// gives global loads with a scalar base.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gptr<T> gp(T* p) {
  return (gptr<T>)p;
}

template <bool NT, typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
  if constexpr (NT) return __builtin_nontemporal_load(gp(p));
  else return *gp(p);
}
template <bool NT, typename T>
__device__ __forceinline__ void st_stream(T* p, T v) {
  if constexpr (NT) __builtin_nontemporal_store(v, gp(p));
  else *gp(p) = v;
}

// ------------------------------------------------------------ complex values
// The reference promotes E, H and the CPML state to complex when an object
// with a complex permittivity is registered (grid.py:353-356,
// objects.py:51-52) and then steps with numpy's complex arithmetic. numpy's
// vectorised loop multiplies complex numbers as
//   re = fma(ar, br, -(ai * bi)),  im = fma(ar, bi, ai * br)
// (checked bit for bit against its results), and promotes a real operand x
// to (x, 0), which makes every real-times-complex product exactly the
// componentwise one; sums and differences are componentwise. cplx<S> does
// the same, so the complex path matches the reference bit for bit as well.
// Division is numpy's for complex values (Smith's method, with its
// zero-divisor branch); a real operand of a division is promoted to (x, 0)
// there as well, which is NOT the componentwise quotient. Absorbing objects
// divide (objects.py:215-221).
template <typename S>
struct cplx {
  S re, im;
  __host__ __device__ cplx() : re(0), im(0) {}
  __host__ __device__ cplx(S r) : re(r), im(0) {}
  __host__ __device__ cplx(S r, S i) : re(r), im(i) {}
  friend __host__ __device__ cplx operator+(cplx a, cplx b) { return cplx(a.re + b.re, a.im + b.im); }
  friend __host__ __device__ cplx operator-(cplx a, cplx b) { return cplx(a.re - b.re, a.im - b.im); }
  friend __host__ __device__ cplx operator-(cplx a) { return cplx(-a.re, -a.im); }
  friend __host__ __device__ cplx operator*(cplx a, cplx b) {
    return cplx(fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re));
  }
  friend __host__ __device__ cplx operator/(cplx a, cplx b) {
    if (fabs(b.re) >= fabs(b.im)) {
      if (b.re == S(0) && b.im == S(0)) return cplx(a.re / fabs(b.re), a.im / fabs(b.re));
      const S rat = b.im / b.re, scl = S(1) / (b.re + b.im * rat);
      return cplx((a.re + a.im * rat) * scl, (a.im - a.re * rat) * scl);
    }
    const S rat = b.re / b.im, scl = S(1) / (b.im + b.re * rat);
    return cplx((a.re * rat + a.im) * scl, (a.im * rat - a.re) * scl);
  }
};

template <bool NT, typename S>
__device__ __forceinline__ cplx<S> ld_stream(const cplx<S>* p) {
  const S* q = reinterpret_cast<const S*>(p);
  return cplx<S>(ld_stream<NT>(q), ld_stream<NT>(q + 1));
}
template <bool NT, typename S>
__device__ __forceinline__ void st_stream(cplx<S>* p, cplx<S> v) {
  S* q = reinterpret_cast<S*>(p);
  st_stream<NT>(q, v.re);
  st_stream<NT>(q + 1, v.im);
}

// ------------------------------------------------------------ field storage
// Element i of a field is component i / cs of cell i % cs (component-major).
// get/put/copy/sample take that element index; get_c/get3/put3 take a cell
// and address its components (the main kernels' accesses).
template <typename S, typename R_>
struct PlainField {
  using R = R_;
  using Sample = S;
  static constexpr bool kRecords = false;  // the marching kernels' record path (CompField) is off
  struct Rec {};
  S* p;
  int64_t cs;
  template <int C>
  __device__ __forceinline__ R get_c(int64_t cell) const { return (R)ld_stream<false>(p + C * cs + cell); }
  template <bool NT>
  __device__ __forceinline__ void get3_s(int64_t cell, R v[3]) const {
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] = (R)ld_stream<NT>(p + c * cs + cell);
  }
  template <bool NT>
  __device__ __forceinline__ void put3_s(int64_t cell, const R v[3]) const {
#pragma unroll
    for (int c = 0; c < 3; ++c) st_stream<NT>(p + c * cs + cell, (S)v[c]);
  }
  __device__ __forceinline__ R get(int64_t i) const { return (R)ld_stream<false>(p + i); }
  template <bool NT>
  __device__ __forceinline__ R get_s(int64_t i) const { return (R)ld_stream<NT>(p + i); }
  template <bool NT>
  __device__ __forceinline__ void put_s(int64_t i, R v) const { st_stream<NT>(p + i, (S)v); }
  __device__ __forceinline__ void put(int64_t i, R v) const { st_stream<false>(p + i, (S)v); }
  __device__ __forceinline__ void copy(int64_t dst, int64_t src) const { st_stream<false>(p + dst, ld_stream<false>(p + src)); }
  __device__ __forceinline__ Sample sample(int64_t i) const { return ld_stream<false>(p + i); }
};

// Compensated storage: a value v is kept as the float h holding the top 24
// significant bits of v rounded to 32 significant bits, plus a byte q with
// the next 8 bits -- a 40-bit floating-point format whose byte is the
// compensation term of the float (what rounding v to the float drops).
// Decoding is exact in double: widen h (its mantissa lands in bits 51..29 of
// the double) and insert q at bits 28..21. Encoding rounds |v| to nearest at
// bit 21 by adding 2^20 to the bit pattern (the carry may ripple into the
// exponent, which is the correct rounding of the magnitude), then splits the
// result. Relative error of a store: at most 2^-32. Values below the float
// range (|v| < 2^-126) keep the float's precision only.
__device__ __forceinline__ double comp_decode(float h, uint32_t q) {
  const double w = (double)h;
  const uint32_t lo32 = (uint32_t)__double2loint(w) | (q << 21);
  return __hiloint2double(__double2hiint(w), (int)lo32);
}
__device__ __forceinline__ float comp_hi(double v, uint32_t& q) {
  const uint64_t u = (uint64_t)__double_as_longlong(v) + (1ull << 20);
  q = (uint32_t)((u >> 21) & 0xffu);
  return (float)__longlong_as_double((long long)(u & ~((1ull << 29) - 1)));
}

// A cell is one 16-byte record, (N, 4) floats: its x, y, z floats and a
// 32-bit word holding the three compensation bytes (byte c for component c),
// so a cell's three components and their compensation come in one 16-byte
// access (four 4-byte accesses with the floats (3, N) and the words apart),
// with no padding. Loads are 16-byte vectors; stores are four float stores
// (merged into one 16-byte store by the backend): a store of the vector type
// may alias any table, with the effect described next. The words are accessed
// as float bit patterns (moved, never computed with): stores of
// an integer type could alias the integer fields of the face and object
// tables that the same loops read, and type-based alias analysis would then
// turn those uniform table reads into vector loads that wait for every
// outstanding store, once per face and cell.
// 16-byte vector of four floats (GCC vector extension: element access by [])
typedef float f32x4 __attribute__((vector_size(16)));

struct CompField {
  using R = double;
  using Sample = float;
  // The marching kernels keep the previous plane's record as loaded and
  // decode it where it is used: a decoded carry is computed right after its
  // load, ahead of the branch that guards a cell's store, so the wave would
  // wait for that load before issuing the plane's others (two memory round
  // trips per plane instead of one).
  static constexpr bool kRecords = true;
  using Rec = f32x4;
  float* rec;  // (N, 4): x, y, z rounded to float, then the compensation word
  int64_t cs;
  __device__ __forceinline__ int comp_of(int64_t i) const { return i >= 2 * cs ? 2 : (i >= cs ? 1 : 0); }
  template <bool NT>
  __device__ __forceinline__ f32x4 record(int64_t cell) const {
    return ld_stream<NT>(reinterpret_cast<const f32x4*>(rec) + cell);
  }
  __device__ __forceinline__ uint32_t word(int64_t cell) const { return __float_as_uint(gp(rec)[4 * cell + 3]); }
  __device__ __forceinline__ void set_byte(int64_t cell, int c, uint32_t q) const {
    gp(rec)[4 * cell + 3] = __uint_as_float((word(cell) & ~(0xffu << (8 * c))) | (q << (8 * c)));
  }
  __device__ __forceinline__ double get(int64_t i) const {
    const int c = comp_of(i);
    const int64_t cell = i - c * cs;
    return comp_decode(gp(rec)[4 * cell + c], (word(cell) >> (8 * c)) & 0xffu);
  }
  __device__ __forceinline__ void put(int64_t i, double v) const {
    const int c = comp_of(i);
    const int64_t cell = i - c * cs;
    uint32_t q;
    gp(rec)[4 * cell + c] = comp_hi(v, q);
    set_byte(cell, c, q);
  }
  __device__ __forceinline__ void copy(int64_t dst, int64_t src) const {
    const int cd = comp_of(dst), cs_ = comp_of(src);
    const int64_t celld = dst - cd * cs, cells = src - cs_ * cs;
    gp(rec)[4 * celld + cd] = gp(rec)[4 * cells + cs_];
    set_byte(celld, cd, (word(cells) >> (8 * cs_)) & 0xffu);
  }
  // detectors sample the value rounded to float
  __device__ __forceinline__ float sample(int64_t i) const { return (float)get(i); }
  __device__ __forceinline__ Rec rec3(int64_t cell) const { return record<false>(cell); }
  template <int C>
  static __device__ __forceinline__ double decode_c(const f32x4& r) {
    return comp_decode(r[C], (__float_as_uint(r[3]) >> (8 * C)) & 0xffu);
  }
  static __device__ __forceinline__ void decode3(const f32x4& r, double v[3]) {
    const uint32_t w = __float_as_uint(r[3]);
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] = comp_decode(r[c], (w >> (8 * c)) & 0xffu);
  }
  template <int C>
  __device__ __forceinline__ double get_c(int64_t cell) const { return decode_c<C>(record<false>(cell)); }
  template <bool NT>
  __device__ __forceinline__ void get3_s(int64_t cell, double v[3]) const {
    decode3(record<NT>(cell), v);
  }
  template <bool NT>
  __device__ __forceinline__ void put3_s(int64_t cell, const double v[3]) const {
    f32x4 r;
    uint32_t w = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      uint32_t q;
      r[c] = comp_hi(v[c], q);
      w |= q << (8 * c);
    }
    r[3] = __uint_as_float(w);
#pragma unroll
    for (int c = 0; c < 4; ++c) st_stream<NT>(rec + 4 * cell + c, (float)r[c]);
  }
};

// --------------------------------------------------------------- materials
// Materials are read in two parts: key(i) -- what can be loaded a plane
// ahead (the palette index) -- and get3(key, i). The marching kernels load
// the next plane's key with the current plane, so a palette lookup does not
// wait for its index.
struct NoKey {};

template <typename S, typename R>
struct MatArray {
  using Key = NoKey;
  const S* p;  // (3, N)
  int64_t cs;
  template <bool NT>
  __device__ __forceinline__ Key key(int64_t) const { return {}; }
  template <bool NT>
  __device__ __forceinline__ void get3(Key, int64_t i, R m[3]) const {
#pragma unroll
    for (int c = 0; c < 3; ++c) m[c] = (R)ld_stream<NT>(p + c * cs + i);
  }
};

template <typename R>
struct MatPalette {
  using Key = int;
  const uint8_t* id;  // (N)
  const double* tab;  // (256, 3)
  template <bool NT>
  __device__ __forceinline__ Key key(int64_t i) const { return (int)ld_stream<NT>(id + i); }
  template <bool NT>
  __device__ __forceinline__ void get3(Key k, int64_t, R m[3]) const {
#pragma unroll
    for (int c = 0; c < 3; ++c) m[c] = (R)gp(tab)[3 * k + c];
  }
};

// ------------------------------------------------------------ CPML helpers
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
  T* const psi = reinterpret_cast<T*>(IS_E ? fd.psiE : fd.psiH);
  const gptr<const double> bt = gp(IS_E ? fd.bE : fd.bH);
  const gptr<const double> ct = gp(IS_E ? fd.cE : fd.cH);
  const T b = (T)bt[l], c = (T)ct[l];
  const T d0 = ax == 0 ? zx : (ax == 1 ? zy : yz);
  const T d1 = ax == 0 ? yx : (ax == 1 ? xy : xz);
  T p0 = ld_stream<false>(psi + off) * b;
  T p1 = ld_stream<false>(psi + fd.slab + off) * b;
  if (IS_E ? (l >= 1) : (l <= fd.t - 2)) {
    p0 = p0 + d0 * c;
    p1 = p1 + d1 * c;
  }
  st_stream<false>(psi + off, p0);
  st_stream<false>(psi + fd.slab + off, p1);
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

// E factor (1 - f) / (1 + f) of an absorbing object (objects.py:215). With
// complex fields the reference divides in the dtype of f: complex (numpy's
// division) when the object's permittivity is complex, real when it is real
// (its arrays stay real on complex fields), and the real quotient then
// multiplies E as (r, 0).
template <typename T>
__device__ __forceinline__ T absorb_ratio(T f, int) {
  return (T(1) - f) / (T(1) + f);
}
template <typename S>
__device__ __forceinline__ cplx<S> absorb_ratio(cplx<S> f, int real) {
  if (real) return cplx<S>((S(1) - f.re) / (S(1) + f.re), S(0));
  return (cplx<S>(1) - f) / (cplx<S>(1) + f);
}

// Row of the anisotropic 3x3 product inv @ curl (objects.py:262-269, numpy's
// einsum "ijk,ikl->ijl"). Real: the association (m0 c0 + m2 c2) + m1 c1
// numpy's einsum uses for this 3-term contraction. Complex: einsum's complex
// loop, products without fused multiply-adds (re = ar br - ai bi,
// im = ar bi + ai br) summed in order; a real inverse on complex fields is
// promoted to (x, 0) by einsum, which the same loop covers.
template <typename T>
__device__ __forceinline__ T aniso_row(T m0, T m1, T m2, T c0, T c1, T c2) {
  return (m0 * c0 + m2 * c2) + m1 * c1;
}
template <typename S>
__device__ __forceinline__ cplx<S> einsum_product(cplx<S> a, cplx<S> b) {
  return cplx<S>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename S>
__device__ __forceinline__ cplx<S> aniso_row(cplx<S> m0, cplx<S> m1, cplx<S> m2, cplx<S> c0, cplx<S> c1,
                                             cplx<S> c2) {
  return (einsum_product(m0, c0) + einsum_product(m1, c1)) + einsum_product(m2, c2);
}

// Object updates of one cell, registration order (objects.py:118-269); the
// object arrays are stored as O and computed in T.
template <typename T, typename O>
__device__ __forceinline__ void objects_update_E(const ObjDev* __restrict__ objs, int nobjs, T sc, int i, int j,
                                                 int k, const T curl[3], T e[3]) {
  for (int o = 0; o < nobjs; ++o) {
    const ObjDev& ob = objs[o];
    if (i < ob.x0 || i >= ob.x1 || j < ob.y0 || j >= ob.y1 || k < ob.z0 || k >= ob.z1) continue;
    const int64_t lo = ((int64_t)(i - ob.x0) * (ob.y1 - ob.y0) + (j - ob.y0)) * (ob.z1 - ob.z0) + (k - ob.z0);
    const O* const inv = reinterpret_cast<const O*>(ob.inv);
    auto at = [](const O* q, int64_t x) { return (T)ld_stream<false>(q + x); };
    if (ob.kind == FDTD_OBJ_PLAIN) {
#pragma unroll
      for (int c = 0; c < 3; ++c) e[c] = e[c] + (sc * at(inv, lo * 3 + c)) * curl[c];
    } else if (ob.kind == FDTD_OBJ_ABSORBING) {
      const O* const fa = reinterpret_cast<const O*>(ob.absorb);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T f = at(fa, lo * 3 + c);
        e[c] = e[c] * absorb_ratio(f, ob.real);
        e[c] = e[c] + ((sc * at(inv, lo * 3 + c)) * curl[c]) / (T(1) + f);
      }
    } else {
      const O* const m = inv + lo * 9;
      T v[3];
#pragma unroll
      for (int r = 0; r < 3; ++r)
        v[r] = aniso_row(at(m, 3 * r), at(m, 3 * r + 1), at(m, 3 * r + 2), curl[0], curl[1], curl[2]);
#pragma unroll
      for (int r = 0; r < 3; ++r) e[r] = e[r] + sc * v[r];
    }
  }
}

// Logical tile of hardware workgroup w: workgroups are dealt round-robin to
// the 8 XCDs (each with its own L2), so consecutive logical tiles -- which
// share halo planes -- are given to the same XCD.
__device__ __forceinline__ unsigned xcd_tile(unsigned w, unsigned nb) {
  const unsigned per = nb / 8, rem = nb % 8, x = w % 8, idx = w / 8;
  return x < rem ? x * (per + 1) + idx : rem * (per + 1) + (x - rem) * per + idx;
}

// Tile of workgroup w in the single-grid kernels of large grids, or -1 for
// an idle workgroup. The x-chunks (`per_chunk` tiles each) are dealt to the
// XCDs cyclically -- chunk c to XCD c mod 8 -- and each XCD walks the tiles of
// its chunks in order, so at any moment the 8 XCDs work on 8 neighbouring
// chunks: the CPML slabs at both x ends, which cost more per cell, are shared
// by all XCDs instead of falling to the first and last (workgroups are
// dispatched in order, so one slow XCD holds back the others). The price is
// the x-halo plane of each chunk, read by another XCD; grids small enough to
// live in the L2s and the memory-side cache keep the contiguous dealing of
// xcd_tile. The grid has 8 * ceil(nchunks / 8) * per_chunk workgroups.
__device__ __forceinline__ int64_t xcd_chunk_tile(unsigned w, unsigned per_chunk, unsigned nchunks) {
  const unsigned x = w % 8, k = w / 8;
  const unsigned tx = (k / per_chunk) * 8 + x;
  return tx < nchunks ? (int64_t)tx * per_chunk + k % per_chunk : -1;
}

// Half-open box of cells {x0, x1, y0, y1, z0, z1}; its cells are numbered
// z fastest.
struct Box {
  int32_t x0, x1, y0, y1, z0, z1;
};

__host__ __device__ __forceinline__ int64_t box_cells(const Box& b) {
  return (int64_t)(b.x1 - b.x0) * (b.y1 - b.y0) * (b.z1 - b.z0);
}

// index of cell (i,j,k) inside box b, or -1
__device__ __forceinline__ int64_t box_index(const Box& b, int i, int j, int k) {
  if (i < b.x0 || i >= b.x1 || j < b.y0 || j >= b.y1 || k < b.z0 || k >= b.z1) return -1;
  return ((int64_t)(i - b.x0) * (b.y1 - b.y0) + (j - b.y0)) * (b.z1 - b.z0) + (k - b.z0);
}

// grid cell of box index p
__device__ __forceinline__ int64_t box_cell(const Box& b, const Dims& d, int64_t p) {
  const int64_t nz = b.z1 - b.z0, ny = b.y1 - b.y0;
  const int64_t k = b.z0 + p % nz, j = b.y0 + (p / nz) % ny, i = b.x0 + p / (nz * ny);
  return cell_of(d, i, j, k);
}

// ------------------------------------------------------------ main kernels
constexpr int FOLDED_POINTS = 4;

template <class FS, class MS>
struct MainArgs {
  using R = typename FS::R;
  Dims d;
  R sc;
  FS F;  // field updated in this half step (E or H)
  FS G;  // field whose curl drives it (H or E)
  MS M;  // inv_eps or inv_mu
  const FaceDev* faces;
  int32_t nfaces;
  const ObjDev* objs;
  int32_t nobjs;
  int32_t bz_shift;  // block extent along z is 1 << bz_shift
  int32_t cx;        // planes per workgroup along x
  int32_t by;        // block extent along y (threads per block = bz * by)
  int32_t cyclic;    // single-grid launch: x-chunks dealt cyclically to the XCDs
  int32_t flat;      // 0: tiles of bz z-cells by `by` z-rows; else tiles of `flat`
                     // (= threads per block) consecutive cells of the (y, z) plane
  // point sources folded into the E kernel (no periodic stage, no detector):
  // their cells, in registration order, and the waveform table
  int32_t npts;
  int64_t pt_cell[FOLDED_POINTS];
  const double* wave;  // row s = source s
  int64_t wave_cap;
  int64_t q;
  // differentiation record (E kernel, recording variant): the curl of this
  // half step inside box `rb`, (3, cells of rb)
  R* rec;
  Box rb;
};

// Tiles of one x-chunk: row tiles (bz z-cells by `by` z-rows) or flat tiles
// (`flat` consecutive cells of the (y, z) plane, wrapping from the end of one
// z-row to the start of the next, so that only the plane's last tile has
// idle lanes whatever nz is).
__host__ __device__ __forceinline__ unsigned plane_tiles(const Dims& d, int bz_shift, int by, int flat) {
  if (flat) return (unsigned)(((int64_t)d.ny * d.nz + flat - 1) / flat);
  const int bz = 1 << bz_shift;
  return (unsigned)(((d.nz + bz - 1) / bz) * ((d.ny + by - 1) / by));
}

// This thread's column (j, k) in tile `tile`, the tile's x range [i0, i1),
// and whether the column is inside the grid: threads past the edge take the
// nearest cell (so that their loads stay in bounds) and store nothing.
template <class A>
__device__ __forceinline__ void tile_column(const A& a, unsigned tile, int& j, int& k, int& i0, int& i1, bool& act) {
  const Dims& d = a.d;
  unsigned tx;
  if (a.flat) {
    const int plane = d.ny * d.nz;  // the host uses flat tiles for planes below 2^31 cells
    const unsigned per = (unsigned)((plane + a.flat - 1) / a.flat);
    const int p0 = (int)(tile % per) * a.flat + (int)threadIdx.x;
    tx = tile / per;
    act = p0 < plane;
    const int p = min(p0, plane - 1);
    j = p / d.nz;
    k = p - j * d.nz;
  } else {
    const int bzs = a.bz_shift, bz = 1 << bzs, by = a.by;
    const int ntz = (d.nz + bz - 1) >> bzs, nty = (d.ny + by - 1) / by;
    const int tz = tile % ntz, ty = (tile / ntz) % nty;
    tx = tile / (ntz * nty);
    const int k0 = tz * bz + (threadIdx.x & (bz - 1)), j0 = ty * by + (threadIdx.x >> bzs);
    act = k0 < d.nz && j0 < d.ny;
    k = min(k0, d.nz - 1);
    j = min(j0, d.ny - 1);
  }
  i0 = (int)tx * a.cx;
  i1 = min(i0 + a.cx, d.nx);
}

// Main E kernel. A workgroup owns a (y, z) tile (tile_column) and marches
// along x over a chunk of `cx` planes: the x-neighbours come from registers
// of the previous plane, the y- and z-neighbours from cache (the rows were
// just streamed in by this or the neighbouring wavefront). Every load of a
// plane is issued unconditionally and at once, so a plane costs one memory
// round trip: threads past the grid's edge read the nearest cell and store
// nothing, and at the low edges the neighbour offset is 0 (the difference is
// dropped).
template <class FS, class MS, typename O, bool OBJS, bool NT, bool REC = false>
__device__ __forceinline__ void main_E_tile(const MainArgs<FS, MS>& __restrict__ a, unsigned tile, int64_t q) {
  using T = typename FS::R;
  const Dims d = a.d;
  int j, k, i0, i1;
  bool act;
  tile_column(a, tile, j, k, i0, i1, act);
  const int64_t sy = d.nz, sx = d.px;
  const int64_t zm = k > 0 ? 1 : 0, ym = j > 0 ? sy : 0;
  const T zero = T(0);
  int64_t cell = (int64_t)i0 * sx + (int64_t)j * sy + k;
  T hz_prev = zero, hy_prev = zero;
  typename FS::Rec hrec{};  // records: H of plane i-1 as loaded
  if (i0 > 0) {
    if constexpr (FS::kRecords) {
      hrec = a.G.rec3(cell - sx);
    } else {
      hz_prev = a.G.template get_c<2>(cell - sx);
      hy_prev = a.G.template get_c<1>(cell - sx);
    }
  }
  typename MS::Key key = a.M.template key<NT>(cell);
  for (int i = i0; i < i1; ++i, cell += sx) {
    T hv[3], m[3], e[3];
    const typename MS::Key kc_ = key;
    if (i + 1 < i1) key = a.M.template key<NT>(cell + sx);  // the next plane's, loaded with this one
    T hy_km, hx_km, hz_jm, hx_jm;
    if constexpr (FS::kRecords) {
      const typename FS::Rec hr = a.G.rec3(cell);
      FS::decode3(hr, hv);
      const f32x4 zr = a.G.rec3(cell - zm), yr = a.G.rec3(cell - ym);
      hx_km = FS::template decode_c<0>(zr), hy_km = FS::template decode_c<1>(zr);
      hx_jm = FS::template decode_c<0>(yr), hz_jm = FS::template decode_c<2>(yr);
      hz_prev = FS::template decode_c<2>(hrec);
      hy_prev = FS::template decode_c<1>(hrec);
      hrec = hr;
    } else {
      a.G.template get3_s<false>(cell, hv);
      hy_km = a.G.template get_c<1>(cell - zm), hx_km = a.G.template get_c<0>(cell - zm);
      hz_jm = a.G.template get_c<2>(cell - ym), hx_jm = a.G.template get_c<0>(cell - ym);
    }
    a.M.template get3<NT>(kc_, cell, m);
    a.F.template get3_s<NT>(cell, e);
    const T hx = hv[0], hy = hv[1], hz = hv[2];
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
    T kc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      kc[c] = a.sc * m[c];
      e[c] = e[c] + kc[c] * curl[c];
    }
    if (act) {
      if (REC) {
        const int64_t r = box_index(a.rb, i, j, k);
        if (r >= 0) {
          const int64_t nr = box_cells(a.rb);
#pragma unroll
          for (int c = 0; c < 3; ++c) a.rec[c * nr + r] = curl[c];
        }
      }
      if (OBJS) objects_update_E<T, O>(a.objs, a.nobjs, a.sc, i, j, k, curl, e);
      if (a.nfaces) pml_cell<true>(a.faces, a.nfaces, d, i, j, k, df, kc, e);
      for (int s = 0; s < a.npts; ++s)  // sources.py:109, registration order
        if (a.pt_cell[s] == cell) e[2] = e[2] + (T)gp(a.wave)[s * a.wave_cap + q];
      a.F.template put3_s<NT>(cell, e);
    }
    if constexpr (!FS::kRecords) {
      hz_prev = hz;
      hy_prev = hy;
    }
  }
}

// tiles per x-chunk and the number of x-chunks of a single-grid launch
template <class FS, class MS>
__device__ __forceinline__ void chunk_counts(const MainArgs<FS, MS>& a, unsigned& per_chunk, unsigned& nchunks) {
  per_chunk = plane_tiles(a.d, a.bz_shift, a.by, a.flat);
  nchunks = (unsigned)((a.d.nx + a.cx - 1) / a.cx);
}

template <class FS, class MS>
__device__ __forceinline__ int64_t launch_tile(const MainArgs<FS, MS>& a, unsigned& ntiles) {
  unsigned per, nch;
  chunk_counts(a, per, nch);
  ntiles = per * nch;
  return a.cyclic ? xcd_chunk_tile(blockIdx.x, per, nch) : (int64_t)xcd_tile(blockIdx.x, gridDim.x);
}

// Occupancy floor of the E kernels: the object variants otherwise take ~118
// VGPRs (4 waves per SIMD) for the anisotropic 3x3 path. Under a floor of 6
// the double-precision ones fit in ~76 VGPRs without spilling, the float32
// ones under a floor of 8 in ~46 (a floor of 8 makes the double ones spill);
// complex128 values need twice the registers and keep a floor of 4. The
// compensated kernels without objects fit 8 waves in 62 VGPRs (66 and 7
// waves unbounded; 0.6 % faster per step at 256^3).
template <class FS, bool OBJS = true>
struct EWaves {
  static constexpr int value =
      FS::kRecords && !OBJS ? 8 : (sizeof(typename FS::R) == 4 ? 8 : (sizeof(typename FS::R) == 8 ? 6 : 4));
};

template <class FS, class MS, typename O, bool OBJS, bool NT, bool REC>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(EWaves<FS, OBJS>::value)))
k_main_E(MainArgs<FS, MS> a) {
  unsigned n;
  const int64_t t = launch_tile(a, n);
  if (t >= 0) main_E_tile<FS, MS, O, OBJS, NT, REC>(a, (unsigned)t, a.q);
}

// Main H kernel: the mirror image of the E kernel. Its differences are
// forward ones, so workgroups take the tiles in decreasing order (the +y/+z
// neighbour tiles were read just before, by the same XCD) and march along -x,
// keeping E of plane i+1 in registers.
template <class FS, class MS, bool NT>
__device__ __forceinline__ void main_H_tile(const MainArgs<FS, MS>& __restrict__ a, unsigned tile) {
  using T = typename FS::R;
  const Dims d = a.d;
  int j, k, i0, i1;
  bool act;
  tile_column(a, tile, j, k, i0, i1, act);
  const int64_t sy = d.nz, sx = d.px;
  const int64_t zp = k < d.nz - 1 ? 1 : 0, yp = j < d.ny - 1 ? sy : 0;
  const T zero = T(0);
  int64_t cell = (int64_t)(i1 - 1) * sx + (int64_t)j * sy + k;
  T ez_next = zero, ey_next = zero;  // E at plane i+1
  typename FS::Rec erec{};          // records: E of plane i+1 as loaded
  if (i1 < d.nx) {
    if constexpr (FS::kRecords) {
      erec = a.G.rec3(cell + sx);
    } else {
      ez_next = a.G.template get_c<2>(cell + sx);
      ey_next = a.G.template get_c<1>(cell + sx);
    }
  }
  typename MS::Key key = a.M.template key<NT>(cell);
  for (int i = i1 - 1; i >= i0; --i, cell -= sx) {
    T ev[3], m[3], h[3];
    const typename MS::Key kc_ = key;
    if (i > i0) key = a.M.template key<NT>(cell - sx);  // the next plane's (marching -x)
    T ey_kp, ex_kp, ez_jp, ex_jp;
    if constexpr (FS::kRecords) {
      const typename FS::Rec er = a.G.rec3(cell);
      FS::decode3(er, ev);
      const f32x4 zr = a.G.rec3(cell + zp), yr = a.G.rec3(cell + yp);
      ex_kp = FS::template decode_c<0>(zr), ey_kp = FS::template decode_c<1>(zr);
      ex_jp = FS::template decode_c<0>(yr), ez_jp = FS::template decode_c<2>(yr);
      ez_next = FS::template decode_c<2>(erec);
      ey_next = FS::template decode_c<1>(erec);
      erec = er;
    } else {
      a.G.template get3_s<false>(cell, ev);
      ey_kp = a.G.template get_c<1>(cell + zp), ex_kp = a.G.template get_c<0>(cell + zp);
      ez_jp = a.G.template get_c<2>(cell + yp), ex_jp = a.G.template get_c<0>(cell + yp);
    }
    a.M.template get3<NT>(kc_, cell, m);
    a.F.template get3_s<NT>(cell, h);
    const T ex = ev[0], ey = ev[1], ez = ev[2];
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
    T kc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      kc[c] = a.sc * m[c];
      h[c] = h[c] - kc[c] * curl[c];
    }
    if (act) {
      // objects leave H unchanged (objects.py:131-137, 223-229, 271-277)
      if (a.nfaces) pml_cell<false>(a.faces, a.nfaces, d, i, j, k, df, kc, h);
      a.F.template put3_s<NT>(cell, h);
    }
    if constexpr (!FS::kRecords) {
      ez_next = ez;
      ey_next = ey;
    }
  }
}

template <class FS, class MS, bool NT>
__global__ void __launch_bounds__(1024) k_main_H(MainArgs<FS, MS> a) {
  unsigned n;
  const int64_t t = launch_tile(a, n);
  if (t >= 0) main_H_tile<FS, MS, NT>(a, n - 1 - (unsigned)t);
}

// ------------------------------------------------------------ batched kernels
// Many independent grids stepped by one launch per half step over a work
// list of positions (first[s] = first position of grid s, item_sim[u] = the
// grid of position u). Workgroup w takes positions w', w' + G, w' + 2G, ...
// (G workgroups). With G equal to the number of positions every workgroup has
// one tile, as for a single grid; a smaller G makes the grid of workgroups
// persistent. Two dealings to the 8 XCDs (workgroup w runs on XCD w mod 8):
//  * contiguous (small batches): the positions are every grid's tiles in
//    turn and w' = w dealt so that each XCD gets a contiguous run of them
//    (neighbouring tiles share halo planes in its L2);
//  * cyclic (batches whose fields outgrow the caches): w' = w, and the
//    x-chunks of all grids, taken in turn, are dealt to the XCDs in rotation
//    -- chunk g of that sequence to XCD g mod 8 (grid s starts at XCD rot[s])
//    -- each XCD walking the tiles of its chunks in order. Every XCD then gets
//    an eighth of every grid, so grids of unequal cost per tile (CPML-heavy
//    small grids beside large ones) load the XCDs evenly, and at any moment
//    the 8 XCDs work on neighbouring chunks, as a single grid's cyclic
//    launch does. A grid spans whole rounds of 8 chunks; the positions of a
//    round that fall before its first chunk or after its last are idle.
// The time index lives in device memory so that a captured graph of steps
// can be replayed: the E kernel reads it for the folded point sources, the H
// kernel advances it.
template <class FS, class MS>
struct BatchArgs {
  const MainArgs<FS, MS>* sims;  // per grid, launch shape included
  const int64_t* first;          // nsim + 1 entries
  const int32_t* item_sim;       // grid of each position
  const int32_t* rot;            // cyclic dealing: XCD of each grid's first chunk
  int64_t total;                 // positions (cyclic: a multiple of 8)
  int64_t* q;                    // time index of the next E half step
  int32_t cyclic;
};

// Tile of grid a at position `local` of its span under cyclic dealing
// (the grid's first chunk on XCD rot), or -1 for an idle position.
template <class FS, class MS>
__device__ __forceinline__ int64_t batch_cyclic_tile(const MainArgs<FS, MS>& a, unsigned local, int rot) {
  unsigned per, nch;
  chunk_counts(a, per, nch);
  const unsigned x = local % 8, k = local / 8;
  const int c = (int)((k / per) * 8 + x) - rot;
  return c >= 0 && c < (int)nch ? (int64_t)c * per + k % per : -1;
}

// One position of a batched E launch. The position and its grid are the
// same for the whole workgroup: say so, so that the grid's arguments are read
// with scalar loads (and, through the restrict reference of the tile body,
// not read again after the field stores). The launch arguments come by value:
// taken by reference, the pointers read from them lost their global address
// space, and the face tables were read with vector (flat) loads, each cell's
// face loop waiting on memory (twice the time per launch).
template <class FS, class MS, typename O, bool OBJS, bool NT>
__device__ __forceinline__ void batch_E_at(BatchArgs<FS, MS> b, int64_t it, int64_t q) {
  const int64_t u = __builtin_amdgcn_readfirstlane((int)it);
  const int s = __builtin_amdgcn_readfirstlane(gp(b.item_sim)[u]);
  int64_t t = u - b.first[s];
  if (b.cyclic) {
    t = batch_cyclic_tile(b.sims[s], (unsigned)t, b.rot[s]);
    if (t < 0) return;
  }
  main_E_tile<FS, MS, O, OBJS, NT>(b.sims[s], (unsigned)t, q);
}

// One position of a batched H launch: each grid's tiles in decreasing order
// (the mirror of the E kernel's order, as for one grid; contiguous dealing
// walks the whole list from its end).
template <class FS, class MS, bool NT>
__device__ __forceinline__ void batch_H_at(BatchArgs<FS, MS> b, int64_t it) {
  const int64_t u = __builtin_amdgcn_readfirstlane((int)(b.cyclic ? it : b.total - 1 - it));
  const int s = __builtin_amdgcn_readfirstlane(gp(b.item_sim)[u]);
  int64_t t = u - b.first[s];
  if (b.cyclic) {
    unsigned per, nch;
    chunk_counts(b.sims[s], per, nch);
    t = batch_cyclic_tile(b.sims[s], (unsigned)t, b.rot[s]);
    if (t < 0) return;
    t = (int64_t)per * nch - 1 - t;
  }
  main_H_tile<FS, MS, NT>(b.sims[s], (unsigned)t);
}

// LOOP: a persistent grid of workgroups walking the list; otherwise one
// position per workgroup, with no loop around the tile body (the loop's
// live values made the compiler spill scalar registers to vector lanes,
// ~50 more vector instructions per wavefront).
template <class FS, class MS, typename O, bool OBJS, bool NT, bool LOOP>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(EWaves<FS>::value)))
k_batch_E(BatchArgs<FS, MS> b) {
  const int64_t q = *b.q;
  const int64_t it0 = b.cyclic ? (int64_t)blockIdx.x : (int64_t)xcd_tile(blockIdx.x, gridDim.x);
  if constexpr (LOOP) {
    for (int64_t it = it0; it < b.total; it += gridDim.x) batch_E_at<FS, MS, O, OBJS, NT>(b, it, q);
  } else if (it0 < b.total) {
    batch_E_at<FS, MS, O, OBJS, NT>(b, it0, q);
  }
}

// H half step; one thread advances the time index (nothing else in this
// kernel reads it, and the E kernel of this step has completed).
template <class FS, class MS, bool NT, bool LOOP>
__global__ void __launch_bounds__(1024) k_batch_H(BatchArgs<FS, MS> b) {
  const int64_t it0 = b.cyclic ? (int64_t)blockIdx.x : (int64_t)xcd_tile(blockIdx.x, gridDim.x);
  if constexpr (LOOP) {
    for (int64_t it = it0; it < b.total; it += gridDim.x) batch_H_at<FS, MS, NT>(b, it);
  } else if (it0 < b.total) {
    batch_H_at<FS, MS, NT>(b, it0);
  }
  // after the work: a branch at the top would keep the compiler from
  // treating the tile loop's values as uniform
  if (blockIdx.x == 0 && threadIdx.x == 0) b.q[0] = b.q[0] + 1;
}

// Sets the batched time index (a kernel rather than a memory copy, so that
// nothing of the host stays referenced after the call).
__global__ void k_set_time(int64_t* q, int64_t v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) q[0] = v;
}

// One CPML face applied on its own (faces registered after a periodic
// boundary): F is the field being updated, G the other one. AX, the face's
// normal axis, is a template parameter so that no index depends on a runtime
// branch over the axis.
template <class FS, class MS, bool IS_E, int AX>
__global__ void __launch_bounds__(256) k_face(Dims d, typename FS::R sc, FS F, FS G, MS M, FaceDev fd) {
  using T = typename FS::R;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= fd.slab) return;
  // slab-local coordinates, the slab laid out like the grid with axis AX of length t
  const int64_t sny = AX == 1 ? fd.t : d.ny, snz = AX == 2 ? fd.t : d.nz;
  const int64_t li = p / (sny * snz), lj = (p / snz) % sny, lk = p % snz;
  const int l = (int)(AX == 0 ? li : (AX == 1 ? lj : lk));
  const int n = AX == 0 ? d.nx : (AX == 1 ? d.ny : d.nz);
  const int g = fd.high ? n - fd.t + l : l;
  const int i = AX == 0 ? g : (int)li, j = AX == 1 ? g : (int)lj, k = AX == 2 ? g : (int)lk;
  const int64_t cell = cell_of(d, i, j, k);
  const int64_t st = AX == 0 ? d.px : (AX == 1 ? (int64_t)d.nz : 1);
  // the other field's components differenced by this face: 3-AX-C0, 3-AX-C1
  constexpr int S0 = 3 - AX - (AX == 0 ? 1 : 0), S1 = 3 - AX - (AX == 2 ? 1 : 2);
  const int64_t o0 = S0 * d.cs + cell, o1 = S1 * d.cs + cell;
  T d0 = T(0), d1 = T(0);
  if (IS_E && l >= 1) {
    d0 = G.get(o0) - G.get(o0 - st);
    d1 = G.get(o1) - G.get(o1 - st);
  }
  if (!IS_E && l <= fd.t - 2) {
    d0 = G.get(o0 + st) - G.get(o0);
    d1 = G.get(o1 + st) - G.get(o1);
  }
  T m[3], kc[3], v[3];
  M.template get3<false>(M.template key<false>(cell), cell, m);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    kc[c] = sc * m[c];
    v[c] = F.get(c * d.cs + cell);
  }
  // pml_face reads the two normal-axis differences
  pml_face<IS_E>(fd, d, i, j, k, d0, d1, d0, d1, d0, d1, kc[0], kc[1], kc[2], v[0], v[1], v[2]);
#pragma unroll
  for (int c = 0; c < 3; ++c) F.put(c * d.cs + cell, v[c]);
}

// Periodic plane copy: to_low ? plane 0 <- plane n-1 : plane n-1 <- plane 0.
template <class FS>
__global__ void __launch_bounds__(256) k_periodic(Dims d, FS F, int axis, int to_low) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t n_plane, src, dst;
  if (axis == 0) {
    n_plane = (int64_t)d.ny * d.nz;
    if (p >= n_plane) return;
    const int64_t last = (int64_t)(d.nx - 1) * d.px + p;
    src = to_low ? last : p;
    dst = to_low ? p : last;
  } else if (axis == 1) {
    n_plane = (int64_t)d.nx * d.nz;
    if (p >= n_plane) return;
    const int64_t i = p / d.nz, k = p % d.nz;
    const int64_t first = cell_of(d, i, 0, k), last = cell_of(d, i, d.ny - 1, k);
    src = to_low ? last : first;
    dst = to_low ? first : last;
  } else {
    n_plane = (int64_t)d.nx * d.ny;
    if (p >= n_plane) return;
    const int64_t first = cell_of(d, p / d.ny, p % d.ny, 0), last = first + d.nz - 1;
    src = to_low ? last : first;
    dst = to_low ? first : last;
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) F.copy(c * d.cs + dst, c * d.cs + src);
}

template <class FS>
struct TailArgs {
  Dims d;
  FS F;  // field of this half step
  int32_t is_H;
  int32_t nsrc, ndet;
  const SrcDev* srcs;
  const DetDev* dets;
  const double* wave;
  int64_t wave_cap;
  int64_t q;
  int64_t counter;  // half steps of this kind already taken by the plan
  int64_t counter_H;  // H half steps taken (the soft sources' current sample)
  int32_t ncur;
  const CurDev* curs;
  // differentiation record (E half step): the field inside box `rb` after
  // the sources, (3, cells of rb), when rec is not null
  typename FS::R* rec;
  Box rb;
};

// Sources in registration order, then detector samples; one workgroup so the
// ordering between sources touching the same cell and the detector reads is
// kept by workgroup barriers.
template <class FS>
__global__ void __launch_bounds__(256) k_tail(TailArgs<FS> a) {
  using T = typename FS::R;
  using Sample = typename FS::Sample;
  const int tid = threadIdx.x;
  for (int s = 0; s < a.nsrc; ++s) {
    const SrcDev& src = a.srcs[s];
    if (a.is_H && src.kind != FDTD_SRC_PLANE) continue;
    const T w = (T)gp(a.wave)[s * a.wave_cap + a.q];
    const int64_t base = (int64_t)(a.is_H ? src.compH : src.compE) * a.d.cs;
    if (src.kind == FDTD_SRC_SOFT) {
      // sources.py:604-631: the voltage v of the waveform, plus R times the
      // current its detector measured after the last H half step
      if (tid == 0) {
        const CurDev& cd = a.curs[src.det];
        const T cur = a.q > 0 ? (T)ld_stream<false>(reinterpret_cast<const Sample*>(cd.buf) +
                                                    (a.counter_H - 1 - cd.slot0) * cd.n)
                              : T(0);
        const T u = src.amplitude > 0 ? w + (T)src.amplitude * cur : w;
        const int64_t cell = base + gp(src.cells)[0];
        a.F.put(cell, a.F.get(cell) + u / (T)src.spacing);
        const int64_t slot = a.counter - src.slot0;
        st_stream<false>(reinterpret_cast<Sample*>(src.vin) + slot, (Sample)w);
        st_stream<false>(reinterpret_cast<Sample*>(src.vout) + slot, (Sample)u);
      }
      __syncthreads();
      continue;
    }
    for (int p = tid; p < src.n; p += blockDim.x) {
      const int64_t cell = base + gp(src.cells)[p];
      if (src.kind == FDTD_SRC_POINT)
        a.F.put(cell, a.F.get(cell) + w);
      else if (src.kind == FDTD_SRC_LINE)
        a.F.put(cell, a.F.get(cell) + (T)gp(src.profile)[p] * w);
      else
        a.F.put(cell, (T)src.amplitude * w);
    }
    __syncthreads();
  }
  if (a.is_H) {
    // detectors.py:388-460, X = 0, Y = 1; numpy wraps index -1 to the last
    for (int t = 0; t < a.ncur; ++t) {
      const CurDev& cd = a.curs[t];
      Sample* const buf = reinterpret_cast<Sample*>(cd.buf) + (a.counter - cd.slot0) * cd.n;
      const T dx = (T)cd.spacing;
      for (int p = tid; p < cd.n; p += blockDim.x) {
        const int px = gp(cd.pts)[p], py = gp(cd.pts)[cd.n + p], pz = gp(cd.pts)[2 * cd.n + p];
        auto H = [&](int c, int x, int y, int z) {
          x = x < 0 ? x + a.d.nx : x;
          y = y < 0 ? y + a.d.ny : y;
          z = z < 0 ? z + a.d.nz : z;
          return a.F.get(c * a.d.cs + cell_of(a.d, x, y, z));
        };
        T v1 = (H(0, px, py - 1, pz) - H(0, px, py, pz)) * dx;
        T v2 = (H(1, px, py, pz) - H(1, px - 1, py, pz)) * dx;
        const T c1 = v1 + v2;
        v1 = (H(0, px, py - 1, pz - 1) - H(0, px, py, pz - 1)) * dx;
        v2 = v2 + (H(1, px, py, pz - 1) - H(1, px - 1, py, pz - 1)) * dx;
        const T c2 = v1 + v2;
        st_stream<false>(buf + p, (Sample)((c1 + c2) / T(2)));
      }
    }
  }
  if (a.rec) {
    const int64_t nr = box_cells(a.rb);
    for (int64_t p = tid; p < 3 * nr; p += blockDim.x) {
      const int c = (int)(p / nr);
      a.rec[p] = a.F.get(c * a.d.cs + box_cell(a.rb, a.d, p - c * nr));
    }
  }
  for (int t = 0; t < a.ndet; ++t) {
    const DetDev& det = a.dets[t];
    const int64_t slot = a.counter - (a.is_H ? det.slot0H : det.slot0E);
    Sample* const buf = reinterpret_cast<Sample*>(a.is_H ? det.bufH : det.bufE) + slot * det.n * 3;
    for (int p = tid; p < det.n * 3; p += blockDim.x) {
      const int c = p % 3;
      st_stream<false>(buf + p, a.F.sample(c * a.d.cs + gp(det.cells)[p / 3]));
    }
  }
}

// Non-finite check: flag[0] becomes nonzero when one of the n values of v
// (stride `skip`: with skip = 4 every 4th value is not a field value but a
// compensation word and is ignored) is a NaN or an infinity. The reference
// has no such check (its only guard is the Courant limit, grid.py:123-127);
// it lets a caller find a diverged run without copying the fields back.
template <typename S>
__global__ void __launch_bounds__(256) k_nonfinite(const S* __restrict__ v, int64_t n, int skip, int* flag) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  bool bad = false;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if (skip && i % skip == skip - 1) continue;
    const S x = v[i];
    bad |= !(x - x == S(0));  // NaN - NaN and inf - inf are NaN
  }
  if (bad) atomicOr(flag, 1);
}

// Plain device copy of n 16-byte words: the bandwidth reference for the
// step kernels (a stream that reads and writes every byte once).
__global__ void __launch_bounds__(256) k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}


// ================================================================ adjoint
// The step is linear in (E, H, psi) for fixed materials, so reverse-mode
// differentiation of an objective through T steps applies the transposes of
// the two half steps in reverse order. The kernels below are those exact
// transposes for plain storage with every CPML face folded into the step and
// objects composed into a per-cell affine E update
//   E' = dcoef * E + gcoef * curl_H(H) + CPML + sources
// (sources add constants and drop out). They march like the main kernels:
// the transpose of the H half step gathers backward neighbours (the shape of
// the E kernel), the transpose of the E half step forward ones (the shape of
// the H kernel).

// One CPML face as a transposed kernel sees it at one launch: the adjoint
// psi of its half step that the launch reads and the one it writes (two
// buffers used alternately, (2, slab) values each like the forward psi), the
// half step's b and c profiles (t values) and the slab size; t = 0 when the
// side has no face. A device table holds one per half step, parity, axis and
// side, so a kernel reads each face's entry with one batch of scalar loads
// and no pointer chasing.
template <typename T>
struct AdjSide {
  const T* pin;
  T* pout;
  const double* b;
  const double* c;
  int64_t slab;
  int32_t t, pad;
};

template <typename T>
struct AdjArgs {
  Dims d;
  T sc;
  T* lamE;               // dJ/dE, (3, N)
  T* lamH;               // dJ/dH, (3, N)
  const T* inv_eps;      // grid materials: the CPML terms use S*inv
  const T* inv_mu;
  const T* gcoef;        // (3, N)
  const T* dcoef;        // (3, N) or null (all ones)
  const AdjSide<T>* side;  // [axis][low/high side] of this launch's half step and parity
  Box db;                // design box: grad += S lamE curl
  const T* curl;         // (3, cells of db) recorded for this step
  double* grad;          // (3, cells of db)
  Box pb;                // objective box: lamE += 2 E
  const T* plane;        // (3, cells of pb) recorded for this step
  int32_t bz_shift, cx, by;  // launch shape, as for the main kernels
  // anisotropic objects: the plan's object table and a (3, N) buffer of
  // Y = dJ/d(curl H) of the E half step, written by k_adj_Y
  const ObjDev* objs;
  int32_t nobjs;
  T* ybuf;
};

// Slab layout constants of a face normal to AX: the slab is the grid with
// axis AX shortened to t; `st` is the slab offset of one step along AX.
template <int AX>
__device__ __forceinline__ int64_t slab_off(const Dims& d, int t, int i, int j, int k, int l) {
  if constexpr (AX == 0) return ((int64_t)l * d.ny + j) * d.nz + k;
  else if constexpr (AX == 1) return ((int64_t)i * t + l) * d.nz + k;
  else return ((int64_t)i * d.ny + j) * t + l;
}
template <int AX>
__device__ __forceinline__ int64_t slab_step(const Dims& d, int t) {
  if constexpr (AX == 0) return (int64_t)d.ny * d.nz;
  else if constexpr (AX == 1) return d.nz;
  else return 1;
}

// The faces normal to AX at one cell of k_adj_H (see there): the face's own
// psi transpose and the difference terms of this cell and of the face cell
// one step back along AX (which holds the Y values yb0, yb1). The face's
// axis is a compile-time constant, so components and signs are too. t_lo and
// t_hi are the slab thicknesses of the two sides (0: no face), read once per
// kernel, so a cell outside every slab reads nothing of the face table.
template <int AX, typename T>
__device__ __forceinline__ void adj_faces_H(const AdjArgs<T>& a, int t_lo, int t_hi, int i, int j, int k, T yc0,
                                           T yc1, T yb0, T yb1, T e[3]) {
  constexpr int C0 = AX == 0 ? 1 : 0, C1 = AX == 2 ? 1 : 2;
  constexpr int S0 = 3 - AX - C0, S1 = 3 - AX - C1;
  // phi_H = psi_{C+1} - psi_{C+2} enters H' with a minus sign
  const T sg0 = AX == (C0 + 1) % 3 ? T(-1) : T(1), sg1 = AX == (C1 + 1) % 3 ? T(-1) : T(1);
  const Dims& d = a.d;
  const int g = AX == 0 ? i : (AX == 1 ? j : k);
  const int n = AX == 0 ? d.nx : (AX == 1 ? d.ny : d.nz);
#pragma nounroll
  for (int side = 0; side < 2; ++side) {
    const int t = side ? t_hi : t_lo;
    const int l = side ? g - (n - t) : g;
    if (t == 0 || l < 0 || l > t) continue;  // neither this cell nor the one back is in the slab
    const AdjSide<T>& fd = a.side[AX * 2 + side];
    const gptr<const T> pin = gp(fd.pin);
    const gptr<T> pout = gp(fd.pout);
    const int64_t off = slab_off<AX>(d, t, i, j, k, l), slab = fd.slab;
    if (l < t) {
      const T w0 = pin[off] + sg0 * yc0;
      const T w1 = pin[slab + off] + sg1 * yc1;
      const T b = (T)gp(fd.b)[l];
      pout[off] = b * w0;
      pout[slab + off] = b * w1;
      if (l <= t - 2) {  // psi'(l) holds c(l) (E(l+1) - E(l))
        const T c = (T)gp(fd.c)[l];
        e[S0] = e[S0] - c * w0;
        e[S1] = e[S1] - c * w1;
      }
    }
    if (l >= 1 && l - 1 <= t - 2) {  // ... and this cell is the E(l+1) of slab cell l-1
      const int64_t ob = off - slab_step<AX>(d, t);
      const T w0 = pin[ob] + sg0 * yb0;
      const T w1 = pin[slab + ob] + sg1 * yb1;
      const T c = (T)gp(fd.c)[l - 1];
      e[S0] = e[S0] + c * w0;
      e[S1] = e[S1] + c * w1;
    }
  }
}

// The faces normal to AX at one cell of k_adj_E (see there): the face's own
// psi transpose and the difference terms of this cell and of the face cell
// one step ahead along AX. lc0, lc1 are lamE of the face's components C0, C1
// at this cell, la0, la1 at the cell ahead (already loaded by the caller).
template <int AX, typename T>
__device__ __forceinline__ void adj_faces_E(const AdjArgs<T>& a, int t_lo, int t_hi, int i, int j, int k, int64_t cell,
                                           T lc0, T lc1, T la0, T la1, T h[3]) {
  constexpr int C0 = AX == 0 ? 1 : 0, C1 = AX == 2 ? 1 : 2;
  constexpr int S0 = 3 - AX - C0, S1 = 3 - AX - C1;
  // phi_E = psi_{C+1} - psi_{C+2} enters E' with a plus sign
  const T sg0 = AX == (C0 + 1) % 3 ? T(1) : T(-1), sg1 = AX == (C1 + 1) % 3 ? T(1) : T(-1);
  const Dims& d = a.d;
  const int64_t cs = d.cs;
  const int g = AX == 0 ? i : (AX == 1 ? j : k);
  const int n = AX == 0 ? d.nx : (AX == 1 ? d.ny : d.nz);
  const int64_t st = AX == 0 ? d.px : (AX == 1 ? (int64_t)d.nz : 1);
  const T sc = a.sc;
  auto K = [&](int c, int64_t x) { return sc * a.inv_eps[c * cs + x]; };
#pragma nounroll
  for (int side = 0; side < 2; ++side) {
    const int t = side ? t_hi : t_lo;
    const int l = side ? g - (n - t) : g;
    if (t == 0 || l < -1 || l >= t) continue;  // neither this cell nor the one ahead is in the slab
    const AdjSide<T>& fd = a.side[AX * 2 + side];
    const gptr<const T> pin = gp(fd.pin);
    const gptr<T> pout = gp(fd.pout);
    const int64_t off = slab_off<AX>(d, t, i, j, k, l), slab = fd.slab;
    if (l >= 0) {
      const T w0 = pin[off] + sg0 * (K(C0, cell) * lc0);
      const T w1 = pin[slab + off] + sg1 * (K(C1, cell) * lc1);
      const T b = (T)gp(fd.b)[l];
      pout[off] = b * w0;
      pout[slab + off] = b * w1;
      if (l >= 1) {  // psi'(l) holds c(l) (H(l) - H(l-1))
        const T c = (T)gp(fd.c)[l];
        h[S0] = h[S0] + c * w0;
        h[S1] = h[S1] + c * w1;
      }
    }
    if (l + 1 >= 1 && l + 1 <= t - 1 && g + 1 < n) {  // ... and the H(l-1) of slab cell l+1
      const int64_t ou = off + slab_step<AX>(d, t);
      const T w0 = pin[ou] + sg0 * (K(C0, cell + st) * la0);
      const T w1 = pin[slab + ou] + sg1 * (K(C1, cell + st) * la1);
      const T c = (T)gp(fd.c)[l + 1];
      h[S0] = h[S0] - c * w0;
      h[S1] = h[S1] - c * w1;
    }
  }
}

// Zeroes lam at n cells (one component; lam already offset to it).
template <typename T>
__global__ void __launch_bounds__(256) k_adj_zero(T* lam, const int64_t* cells, int n) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) lam[cells[p]] = T(0);
}

// Transpose of the H half step (grid.py:301-325, boundaries.py:421-431 and
// 461-487). Forward, per face normal to ax and psi component q (H component
// C_q, E component s_q = 3-ax-C_q differenced along ax):
//   psiH'_q = bH psiH_q + cH(l) (E_s(l+1) - E_s(l))     for l <= t-2
//   H'      = H - K_H curl_E(E) + sg_q K_H psiH'_q,       K_H = S inv_mu
// (phi_H = psi_{C+1} - psi_{C+2} enters H' with a minus sign: sg_q = -1 when
// ax = C_q + 1 mod 3, +1 otherwise). With Y = K_H lamH (lamH = dJ/dH' =
// dJ/dH) and w_q = lamPsiH'_q + sg_q Y_{C_q}:
//   lamPsiH_q = bH w_q
//   lamE      = dcoef lamE - curl_E^T(Y) - cH(l) w_q [at E_s(l)]
//                                        + cH(l) w_q [at E_s(l+1)]
// curl_E^T gathers backward neighbours with the masks of curl_E
// (grid.py:39-49): (curl_E^T Y)_x = ([k>=1] Y_y(k-1) - [k<=nz-2] Y_y(k))
// - ([j>=1] Y_z(j-1) - [j<=ny-2] Y_z(j)), and cyclically. The factor dcoef
// comes from the transpose of the following E half step (k_adj_E leaves it);
// after the update the objective's 2 E is added inside the objective box.
// FULL: the tiles cover the (y, z) plane exactly (no thread past an edge), so
// the stores need no guard; with a guard, the branch's untaken side leaves
// the plane's loads unconsumed, and the wait inserted for the next plane's
// register reuse then also waits for this plane's stores.
template <typename T, bool FULL>
__global__ void __launch_bounds__(1024) k_adj_H(AdjArgs<T> a) {
  const Dims d = a.d;
  const int bzs = a.bz_shift, bz = 1 << bzs, by = a.by;
  const int ntz = (d.nz + bz - 1) >> bzs, nty = (d.ny + by - 1) / by;
  const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
  const int tz = tile % ntz, ty = (tile / ntz) % nty, tx = tile / (ntz * nty);
  // as in the main kernels: every load of a plane at once, threads past the
  // edge on the nearest cell (storing nothing), neighbour offset 0 at an edge
  const int k0 = tz * bz + (threadIdx.x & (bz - 1)), j0 = ty * by + (threadIdx.x >> bzs);
  const bool act = FULL || (k0 < d.nz && j0 < d.ny);
  const int k = min(k0, d.nz - 1), j = min(j0, d.ny - 1);
  const int i0 = tx * a.cx, i1 = min(i0 + a.cx, d.nx);
  const int64_t sy = d.nz, sx = d.px, cs = d.cs;
  const int64_t zm = k > 0 ? 1 : 0, ym = j > 0 ? sy : 0;
  const T zero = T(0), sc = a.sc;
  auto Y = [&](int c, int64_t x) { return (sc * a.inv_mu[c * cs + x]) * a.lamH[c * cs + x]; };
  int ts[6];  // slab thickness of each side's face (0: none), read once
#pragma unroll
  for (int q = 0; q < 6; ++q) ts[q] = a.side[q].t;
  int64_t cell = (int64_t)i0 * sx + (int64_t)j * sy + k;
  // inv_mu and lamH of components 1 and 2 at plane i-1, as loaded: Y is formed
  // where it is used. A carried product is computed right after its loads,
  // ahead of the branch that guards the stores, and the wave then waits for
  // them before it issues the plane's other loads (two round trips a plane).
  T mz_prev = zero, lz_prev = zero, my_prev = zero, ly_prev = zero;
  if (i0 > 0) {
    mz_prev = a.inv_mu[2 * cs + cell - sx];
    lz_prev = a.lamH[2 * cs + cell - sx];
    my_prev = a.inv_mu[cs + cell - sx];
    ly_prev = a.lamH[cs + cell - sx];
  }
  for (int i = i0; i < i1; ++i, cell += sx) {
    const T m1 = a.inv_mu[cs + cell], m2 = a.inv_mu[2 * cs + cell];
    const T lh1 = a.lamH[cs + cell], lh2 = a.lamH[2 * cs + cell];
    const T y0 = Y(0, cell), y1 = (sc * m1) * lh1, y2 = (sc * m2) * lh2;
    const T yz_prev = (sc * mz_prev) * lz_prev, yy_prev = (sc * my_prev) * ly_prev;  // Y at plane i-1
    const T yy_km = Y(1, cell - zm), yx_km = Y(0, cell - zm);
    const T yz_jm = Y(2, cell - ym), yx_jm = Y(0, cell - ym);
    T e0 = a.lamE[cell], e1 = a.lamE[cs + cell], e2 = a.lamE[2 * cs + cell];
    if (a.dcoef) {
      e0 = a.dcoef[cell] * e0;
      e1 = a.dcoef[cs + cell] * e1;
      e2 = a.dcoef[2 * cs + cell] * e2;
    }
    const bool km = k >= 1, kp = k <= d.nz - 2, jm = j >= 1, jp = j <= d.ny - 2, im = i >= 1, ip = i <= d.nx - 2;
    const T t0 = ((km ? yy_km : zero) - (kp ? y1 : zero)) - ((jm ? yz_jm : zero) - (jp ? y2 : zero));
    const T t1 = ((im ? yz_prev : zero) - (ip ? y2 : zero)) - ((km ? yx_km : zero) - (kp ? y0 : zero));
    const T t2 = ((jm ? yx_jm : zero) - (jp ? y0 : zero)) - ((im ? yy_prev : zero) - (ip ? y1 : zero));
    e0 = e0 - t0;
    e1 = e1 - t1;
    e2 = e2 - t2;
    if (act) {
      {
        T e[3] = {e0, e1, e2};
        adj_faces_H<0>(a, ts[0], ts[1], i, j, k, y1, y2, yy_prev, yz_prev, e);
        adj_faces_H<1>(a, ts[2], ts[3], i, j, k, y0, y2, yx_jm, yz_jm, e);
        adj_faces_H<2>(a, ts[4], ts[5], i, j, k, y0, y1, yx_km, yy_km, e);
        e0 = e[0];
        e1 = e[1];
        e2 = e[2];
      }
      const int64_t pp = box_index(a.pb, i, j, k);
      if (pp >= 0) {
        const int64_t np = box_cells(a.pb);
        e0 = e0 + T(2) * a.plane[pp];
        e1 = e1 + T(2) * a.plane[np + pp];
        e2 = e2 + T(2) * a.plane[2 * np + pp];
      }
      a.lamE[cell] = e0;
      a.lamE[cs + cell] = e1;
      a.lamE[2 * cs + cell] = e2;
    }
    mz_prev = m2;
    lz_prev = lh2;
    my_prev = m1;
    ly_prev = lh1;
  }
}

// Transpose of the E half step (grid.py:275-299, objects.py:118-221,
// boundaries.py:409-459). Forward, per face normal to ax and psi component q
// (E component C_q, H component s_q = 3-ax-C_q differenced along ax):
//   psiE'_q = bE psiE_q + cE(l) (H_s(l) - H_s(l-1))      for l >= 1
//   E'      = dcoef E + gcoef curl_H(H) + sg_q K_E psiE'_q,  K_E = S inv_eps
// (phi_E enters E' with a plus sign: sg_q = +1 when ax = C_q + 1 mod 3, -1
// otherwise; K_E is the grid's, zero inside objects). With Y = gcoef lamE
// (lamE = dJ/dE') and w_q = lamPsiE'_q + sg_q K_E lamE_{C_q}:
//   lamPsiE_q = bE w_q
//   lamH     += curl_H^T(Y) + cE(l) w_q [at H_s(l)] - cE(l) w_q [at H_s(l-1)]
//   lamE      = dcoef lamE   (left to the next k_adj_H)
//   grad     += S lamE curl  (dE'/d inv = S curl inside the design box)
// curl_H^T gathers forward neighbours with the masks of curl_H
// (grid.py:65-73): (curl_H^T Y)_x = ([k>=1] Y_y - [k<=nz-2] Y_y(k+1))
// - ([j>=1] Y_z - [j<=ny-2] Y_z(j+1)), and cyclically.
// Y = dJ/d(curl H) of the E half step at every cell when the plan has
// anisotropic objects: gcoef lamE (the diagonal part composed by the host)
// plus, inside each anisotropic object, the transpose of its update
// E += S (inv @ curl) (objects.py:254-269): S inv^T lamE. Overlapping objects
// add their terms in the forward step, so their transposes add here too.
template <typename T>
__global__ void __launch_bounds__(256) k_adj_Y(AdjArgs<T> a) {
  const Dims d = a.d;
  const int64_t cs = d.cs;
  const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= cs) return;
  const int64_t r = cell % d.px;  // cell of the x plane; past ny*nz: the plane's padding
  if (r >= (int64_t)d.ny * d.nz) return;
  const T l0 = a.lamE[cell], l1 = a.lamE[cs + cell], l2 = a.lamE[2 * cs + cell];
  T y0 = a.gcoef[cell] * l0, y1 = a.gcoef[cs + cell] * l1, y2 = a.gcoef[2 * cs + cell] * l2;
  const int k = (int)(r % d.nz), j = (int)(r / d.nz), i = (int)(cell / d.px);
  for (int o = 0; o < a.nobjs; ++o) {
    const ObjDev& ob = a.objs[o];
    if (ob.kind != FDTD_OBJ_ANISOTROPIC) continue;
    const Box b{ob.x0, ob.x1, ob.y0, ob.y1, ob.z0, ob.z1};
    const int64_t lo = box_index(b, i, j, k);
    if (lo < 0) continue;
    const gptr<const T> m = gp(reinterpret_cast<const T*>(ob.inv) + lo * 9);
    y0 = y0 + a.sc * ((m[0] * l0 + m[3] * l1) + m[6] * l2);
    y1 = y1 + a.sc * ((m[1] * l0 + m[4] * l1) + m[7] * l2);
    y2 = y2 + a.sc * ((m[2] * l0 + m[5] * l1) + m[8] * l2);
  }
  a.ybuf[cell] = y0;
  a.ybuf[cs + cell] = y1;
  a.ybuf[2 * cs + cell] = y2;
}

// Transpose of a periodic plane copy F[dst] = F[src] along `axis`
// (boundaries.py:184-219; to_low: dst is plane 0, src plane n-1, as for E;
// else the reverse, as for H): lam[src] += lam[dst], lam[dst] = 0.
template <typename T>
__global__ void __launch_bounds__(256) k_adj_periodic(T* lam, Dims d, int axis, int to_low) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t lo_cell, hi_cell;
  if (axis == 0) {
    const int64_t n_plane = (int64_t)d.ny * d.nz;
    if (p >= n_plane) return;
    lo_cell = p;
    hi_cell = (int64_t)(d.nx - 1) * d.px + p;
  } else if (axis == 1) {
    if (p >= (int64_t)d.nx * d.nz) return;
    const int64_t i = p / d.nz, k = p % d.nz;
    lo_cell = cell_of(d, i, 0, k);
    hi_cell = cell_of(d, i, d.ny - 1, k);
  } else {
    if (p >= (int64_t)d.nx * d.ny) return;
    lo_cell = cell_of(d, p / d.ny, p % d.ny, 0);
    hi_cell = lo_cell + d.nz - 1;
  }
  const int64_t dst = to_low ? lo_cell : hi_cell, src = to_low ? hi_cell : lo_cell;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    lam[c * d.cs + src] = lam[c * d.cs + src] + lam[c * d.cs + dst];
    lam[c * d.cs + dst] = T(0);
  }
}

// YB: Y read from ybuf (k_adj_Y ran first) instead of gcoef lamE. (A FULL
// form as for k_adj_H is not used here: without the guard the compiler's
// register reuse made the float kernel wait for the last plane's stores.)
template <typename T, bool YB>
__global__ void __launch_bounds__(1024) k_adj_E(AdjArgs<T> a) {
  const Dims d = a.d;
  const int bzs = a.bz_shift, bz = 1 << bzs, by = a.by;
  const int ntz = (d.nz + bz - 1) >> bzs, nty = (d.ny + by - 1) / by;
  const unsigned tile = gridDim.x - 1 - xcd_tile(blockIdx.x, gridDim.x);
  const int tz = tile % ntz, ty = (tile / ntz) % nty, tx = tile / (ntz * nty);
  const int k0 = tz * bz + (threadIdx.x & (bz - 1)), j0 = ty * by + (threadIdx.x >> bzs);
  const bool act = k0 < d.nz && j0 < d.ny;
  const int k = min(k0, d.nz - 1), j = min(j0, d.ny - 1);
  const int i0 = tx * a.cx, i1 = min(i0 + a.cx, d.nx);
  const int64_t sy = d.nz, sx = d.px, cs = d.cs;
  const int64_t zp = k < d.nz - 1 ? 1 : 0, yp = j < d.ny - 1 ? sy : 0;
  const T zero = T(0), sc = a.sc;
  auto L = [&](int c, int64_t x) { return a.lamE[c * cs + x]; };
  // Y of component c at cell x whose lamE is l
  auto Yc = [&](int c, int64_t x, T l) { return YB ? a.ybuf[c * cs + x] : a.gcoef[c * cs + x] * l; };
  int ts[6];  // slab thickness of each side's face (0: none), read once
#pragma unroll
  for (int q = 0; q < 6; ++q) ts[q] = a.side[q].t;
  int64_t cell = (int64_t)(i1 - 1) * sx + (int64_t)j * sy + k;
  // plane i+1, as loaded: lamE of components 1 and 2 and either their Y
  // (ybuf) or their gcoef (Y = gcoef lamE is formed where it is used, for the
  // reason given in k_adj_H)
  T lz_next = zero, ly_next = zero, fz_next = zero, fy_next = zero;
  if (i1 < d.nx) {
    lz_next = L(2, cell + sx);
    ly_next = L(1, cell + sx);
    fz_next = YB ? a.ybuf[2 * cs + cell + sx] : a.gcoef[2 * cs + cell + sx];
    fy_next = YB ? a.ybuf[cs + cell + sx] : a.gcoef[cs + cell + sx];
  }
  for (int i = i1 - 1; i >= i0; --i, cell -= sx) {
    const T le0 = L(0, cell), le1 = L(1, cell), le2 = L(2, cell);
    const T lk0 = L(0, cell + zp), lk1 = L(1, cell + zp);
    const T lj0 = L(0, cell + yp), lj2 = L(2, cell + yp);
    const T f1 = YB ? a.ybuf[cs + cell] : a.gcoef[cs + cell], f2 = YB ? a.ybuf[2 * cs + cell] : a.gcoef[2 * cs + cell];
    const T y0 = Yc(0, cell, le0), y1 = YB ? f1 : f1 * le1, y2 = YB ? f2 : f2 * le2;
    const T yz_next = YB ? fz_next : fz_next * lz_next, yy_next = YB ? fy_next : fy_next * ly_next;  // Y at plane i+1
    const T yy_kp = Yc(1, cell + zp, lk1), yx_kp = Yc(0, cell + zp, lk0);
    const T yz_jp = Yc(2, cell + yp, lj2), yx_jp = Yc(0, cell + yp, lj0);
    const T l0 = a.lamH[cell], l1 = a.lamH[cs + cell], l2 = a.lamH[2 * cs + cell];
    const bool km = k >= 1, kp = k <= d.nz - 2, jm = j >= 1, jp = j <= d.ny - 2, im = i >= 1, ip = i <= d.nx - 2;
    const T t0 = ((km ? y1 : zero) - (kp ? yy_kp : zero)) - ((jm ? y2 : zero) - (jp ? yz_jp : zero));
    const T t1 = ((im ? y2 : zero) - (ip ? yz_next : zero)) - ((km ? y0 : zero) - (kp ? yx_kp : zero));
    const T t2 = ((jm ? y0 : zero) - (jp ? yx_jp : zero)) - ((im ? y1 : zero) - (ip ? yy_next : zero));
    T h0 = l0 + t0, h1 = l1 + t1, h2 = l2 + t2;
    if (act) {
      {
        T h[3] = {h0, h1, h2};
        adj_faces_E<0>(a, ts[0], ts[1], i, j, k, cell, le1, le2, ly_next, lz_next, h);
        adj_faces_E<1>(a, ts[2], ts[3], i, j, k, cell, le0, le2, lj0, lj2, h);
        adj_faces_E<2>(a, ts[4], ts[5], i, j, k, cell, le0, le1, lk0, lk1, h);
        h0 = h[0];
        h1 = h[1];
        h2 = h[2];
      }
      a.lamH[cell] = h0;
      a.lamH[cs + cell] = h1;
      a.lamH[2 * cs + cell] = h2;
      const int64_t pd = box_index(a.db, i, j, k);
      if (pd >= 0) {
        const int64_t nd = box_cells(a.db);
        const T le[3] = {le0, le1, le2};
#pragma unroll
        for (int c = 0; c < 3; ++c)
          a.grad[c * nd + pd] += (double)sc * (double)le[c] * (double)a.curl[c * nd + pd];
      }
    }
    fz_next = f2;
    fy_next = f1;
    lz_next = le2;
    ly_next = le1;
  }
}

}  // namespace fdtd_dev
