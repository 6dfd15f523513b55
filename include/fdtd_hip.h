/*
 * fdtd_hip.h -- C ABI of libfdtd_hip.so, the MI355X (gfx950) FDTD step engine.
 *
 * The engine replaces the per-statement array work of the reference's
 * Grid.update_E / Grid.update_H (fdtd/grid.py:275-325) and of the component
 * hooks those call, with fused HIP kernels. A plan is built the way the
 * reference registers components (grid[...] = component, fdtd/grid.py:358-375):
 * create it for the grid arrays, then add boundaries, objects, sources and
 * detectors in registration order, then finalize. Stepping is asynchronous on
 * the caller's HIP stream; nothing synchronises the host per step.
 *
 * Conventions
 *   - All array arguments are device pointers borrowed from the caller (the
 *     Python layer keeps them alive as torch tensors); index lists and
 *     coefficient profiles marked "host" are copied at call time.
 *   - Fields E, H and the grid materials are stored component-major:
 *     (3, Nx, Ny, Nz), z fastest. Object arrays keep the reference layout
 *     (nx, ny, nz, 3) or (nx, ny, nz, 3, 3).
 *   - dtype: FDTD_F64 or FDTD_F32 for every floating array of the plan, or
 *     FDTD_F32C: E and H stored as float32 plus an 8-bit compensation term per
 *     value, one 16-byte record per cell (fdtd_plan_set_compensation),
 *     arithmetic in float64, materials, object arrays and CPML state in
 *     float64, detector traces in float32.
 *   - FDTD_C128 / FDTD_C64: the complex-field path of the reference
 *     (fdtd/grid.py:353-356): E, H, CPML state and object inverse
 *     permittivities complex (interleaved re, im; complex128 / complex64),
 *     grid materials real (float64 / float32), detector traces complex;
 *     object arrays complex, holding (x, 0) for an object whose arrays are
 *     real in the reference (FDTD_OBJ_REAL_ARRAYS). Soft sources and current
 *     detectors are refused on this path; so are the material palette,
 *     batching and the adjoint.
 *   - Every function returns 0 on success and a negative code on error;
 *     fdtd_last_error() then describes the failure. No C++ exception crosses
 *     the ABI.
 *   - A plan is used from one host thread at a time.
 */
#ifndef FDTD_HIP_H
#define FDTD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct FdtdPlan FdtdPlan;

enum { FDTD_F64 = 0, FDTD_F32 = 1, FDTD_F32C = 2, FDTD_C128 = 3, FDTD_C64 = 4 };
enum { FDTD_OBJ_PLAIN = 0, FDTD_OBJ_ABSORBING = 1, FDTD_OBJ_ANISOTROPIC = 2 };
/* Flag or-ed into an object kind: the reference holds this object's arrays
 * as real ones (a real permittivity and conductivity). It matters on the
 * complex-field path only, where an absorbing object's factor
 * (1 - f) / (1 + f) is then a real division (fdtd/objects.py:215). */
enum { FDTD_OBJ_REAL_ARRAYS = 16 };
enum { FDTD_SRC_POINT = 0, FDTD_SRC_LINE = 1, FDTD_SRC_PLANE = 2, FDTD_SRC_SOFT = 3 };

/* Version string of the library build. */
const char* fdtd_version(void);

/* Text of the last error raised on the calling thread ("" if none). */
const char* fdtd_last_error(void);

/* New plan over a grid of nx*ny*nz cells.
 * Replaces the state set up by Grid.__init__ (fdtd/grid.py:90-169):
 * E, H (3,nx,ny,nz) read-write (FDTD_F32C: (nx,ny,nz,4) records, see
 * fdtd_plan_set_compensation); inv_eps, inv_mu (3,nx,ny,nz) read-only,
 * read every step so that writes to the grid materials take effect. The x
 * planes of these arrays are ny*nz cells apart unless fdtd_plan_set_pitch
 * says otherwise. */
int fdtd_plan_create(int dtype, int nx, int ny, int nz, double courant_number,
                     void* E, void* H, const void* inv_eps, const void* inv_mu,
                     FdtdPlan** out);

int fdtd_plan_destroy(FdtdPlan* plan);

/* Optional, right after fdtd_plan_create: the x-plane pitch px >= ny*nz of
 * every per-cell array of the plan (fields, materials, records, palette
 * indices, the adjoint's field-shaped buffers): cell (i, j, k) of a component
 * is at i*px + j*nz + k and components are nx*px apart (default px = ny*nz,
 * the reference's contiguous (3, nx, ny, nz) layout, fdtd/grid.py:90-169).
 * A pitch that rounds a plane up to a multiple of 32 cells starts every x
 * plane of every array on a 128-byte line; the padding cells are never read
 * or written. No counterpart in the reference: its arrays are contiguous. */
int fdtd_plan_set_pitch(FdtdPlan* plan, int64_t px);

/* FDTD_F32C only, required. E and H of such a plan are (nx, ny, nz, 4)
 * float32 records, z fastest: x, y, z of the field at the cell, then a uint32
 * word whose byte c holds component c's compensation term (E_lo / H_lo: the
 * word of the first cell, i.e. E + 3 floats / H + 3 floats; other addresses
 * are refused). A value is the float (its top 24 significant bits) extended by
 * the 8 mantissa bits that follow, held in its byte: 32 significant bits in
 * all; a zero byte leaves the float's value unchanged. */
int fdtd_plan_set_compensation(FdtdPlan* plan, void* E_lo, void* H_lo);

/* Optional, any dtype: read the grid materials through a palette instead of
 * the (3, nx, ny, nz) arrays. eps_id / mu_id: uint8 (nx, ny, nz) indices into
 * float64 (256, 3) tables of inverse permittivity / permeability triples (all
 * four device pointers, or all NULL to read the arrays again). The caller
 * rebuilds the palette when the grid materials change. */
int fdtd_plan_set_material_palette(FdtdPlan* plan, const uint8_t* eps_id, const double* eps_table,
                                   const uint8_t* mu_id, const double* mu_table);

/* One CPML face (fdtd/boundaries.py:225-487 with the face subclasses
 * :490-625). axis 0/1/2, high 0 (low face) or 1 (high face), thickness t.
 * bE, cE, bH, cH: host float64 profiles of length t along the normal axis
 * (the only non-zero CPML component). psiE, psiH: device state of the face,
 * 2 components x slab cells, slab laid out like the grid with the normal
 * axis shortened to t, in the plan dtype. The two components are the two
 * field components other than `axis`, in increasing order. */
int fdtd_plan_add_pml(FdtdPlan* plan, int axis, int high, int thickness,
                      const double* bE, const double* cE,
                      const double* bH, const double* cH,
                      void* psiE, void* psiH);

/* Periodic boundary along `axis`: E[0] = E[-1] after the E update and
 * H[-1] = H[0] after the H update (fdtd/boundaries.py:143-219). */
int fdtd_plan_add_periodic(FdtdPlan* plan, int axis);

/* Object over the box [x0,x1) x [y0,y1) x [z0,z1) (fdtd/objects.py:24-277).
 * inv: the object's inverse permittivity, (nx,ny,nz,3) for PLAIN and
 * ABSORBING, (nx,ny,nz,3,3) for ANISOTROPIC. absorption: (nx,ny,nz,3)
 * absorption factors for ABSORBING (objects.py:198-205), else NULL. kind may
 * carry FDTD_OBJ_REAL_ARRAYS. */
int fdtd_plan_add_object(FdtdPlan* plan, int kind, int x0, int x1, int y0, int y1,
                         int z0, int z1, const void* inv, const void* absorption);

/* Source over n cells (fdtd/sources.py:25-501). cells: host linear cell
 * indices ((i*ny + j)*nz + k). POINT adds the waveform value to E[comp_E];
 * LINE adds profile[p] * waveform to E[comp_E]; PLANE assigns
 * amplitude * waveform to E[comp_E] and, in the H half step, to H[comp_H].
 * profile: host float64[n] for LINE, else NULL. The waveform of source s at
 * step q is read from the table set by fdtd_plan_set_waveforms. */
int fdtd_plan_add_source(FdtdPlan* plan, int kind, int n, const int64_t* cells,
                         const double* profile, double amplitude, int comp_E, int comp_H);

/* Device table of float64 waveform values, row s (source registration order)
 * holding q = 0 .. capacity-1 (fdtd/sources.py:93-109, 278-297, 476-486). */
int fdtd_plan_set_waveforms(FdtdPlan* plan, const double* table, int64_t capacity);

/* Detector sampling the 3 field components at n cells after every E half step
 * and every H half step (fdtd/detectors.py:20-280). bufE, bufH: device
 * (capacity, n, 3) arrays in the plan dtype; count_E / count_H samples are
 * already stored in them. */
int fdtd_plan_add_detector(FdtdPlan* plan, int n, const int64_t* cells,
                           void* bufE, void* bufH, int64_t capacity,
                           int64_t count_E, int64_t count_H);

/* Current detector (fdtd/detectors.py:284-476): after every H half step, the
 * loop current around z at each of n points (host arrays px, py, pz; index -1
 * wraps to the last plane as in numpy), in the reference's operation order
 * including its accumulation of the second loop's y-term
 * (detectors.py:447-460), times `spacing`, stored to bufI[slot * n + p]
 * (plan dtype; float for FDTD_F32C), slot = samples so far (`count` at
 * registration). Current detectors are numbered in registration order. */
int fdtd_plan_add_current_detector(FdtdPlan* plan, int n, const int32_t* px, const int32_t* py, const int32_t* pz,
                                   double spacing, void* bufI, int64_t capacity, int64_t count);

/* Resistive soft source (fdtd/sources.py:504-637) at one cell, a source in
 * registration order like the others: in the E half step at time index q,
 * v = waveform table row of this source at q, I = the last sample of current
 * detector `detector` (0 when q == 0), u = v + impedance * I if impedance > 0
 * else v; E_z += u / spacing; v and u are stored to vin / vout[slot] (plan
 * dtype; float for FDTD_F32C), slot = E half steps so far (`count`). */
int fdtd_plan_add_soft_source(FdtdPlan* plan, int64_t cell, double impedance, double spacing, int detector,
                              void* vin, void* vout, int64_t capacity, int64_t count);

/* Samples stored so far by current detector `index` / soft source `index`
 * (numbered among their kind). */
int fdtd_plan_current_detector_count(const FdtdPlan* plan, int index, int64_t* count);
int fdtd_plan_soft_source_count(const FdtdPlan* plan, int index, int64_t* count);

/* Upload the plan tables; required before stepping. */
int fdtd_plan_finalize(FdtdPlan* plan);

/* Samples stored so far by detector `index`. */
int fdtd_plan_detector_counts(const FdtdPlan* plan, int index, int64_t* count_E, int64_t* count_H);

/* One E half step at time index q: everything Grid.update_E does
 * (fdtd/grid.py:275-299). stream: hipStream_t (0 = default stream). */
int fdtd_half_step_E(FdtdPlan* plan, int64_t q, void* stream);

/* One H half step at time index q (fdtd/grid.py:301-325). */
int fdtd_half_step_H(FdtdPlan* plan, int64_t q, void* stream);

/* nsteps full steps starting at time index q0 (fdtd/grid.py:250-273):
 * E half step then H half step at q, then q+1. */
int fdtd_run(FdtdPlan* plan, int64_t q0, int64_t nsteps, void* stream);

/* No counterpart in the reference, whose only guard against divergence is
 * the Courant limit at construction (fdtd/grid.py:123-127): launch on
 * `stream` a check of every E and H value of the plan and OR 1 into the
 * device int *flag (zeroed by the caller) if one is a NaN or an infinity.
 * The plan need not be finalized. */
int fdtd_check_finite(const FdtdPlan* plan, int32_t* flag, void* stream);

/* ---- Batches of grids ----------------------------------------------------
 * No counterpart in the reference (one Grid steps at a time there,
 * fdtd/grid.py:250-273): many independent grids, each a finalized plan with
 * PML faces, objects and at most 4 point sources (no periodic boundaries, no
 * detectors), all of one storage type, stepped together by one launch per
 * half step. The grids' tiles form one work list; the time index lives on
 * the device, so a captured graph of steps can be replayed. */
typedef struct FdtdBatch FdtdBatch;

/* Batch over n plans (borrowed: they must outlive the batch and stay
 * unchanged). */
int fdtd_batch_create(int n, FdtdPlan* const* plans, FdtdBatch** out);

int fdtd_batch_destroy(FdtdBatch* batch);

/* Launch shape: threads per workgroup (256, 512 or 1024), x planes per work
 * item, and the number of workgroups (0, or at least the number of work
 * items: one workgroup per item; fewer: a persistent grid of workgroups that
 * walks the list). Discards a captured graph. */
int fdtd_batch_set_launch(FdtdBatch* batch, int block_threads, int cx, int workgroups);

/* Work items (tiles over all grids) of the current launch shape. */
int fdtd_batch_work_items(const FdtdBatch* batch, int64_t* items);

/* nsteps steps of every grid from time index q0, two launches per step. */
int fdtd_batch_run(FdtdBatch* batch, int64_t q0, int64_t nsteps, void* stream);

/* The same through a HIP graph of steps_per_graph steps, captured on the
 * first call (and again when steps_per_graph changes) and replayed on a
 * stream of the batch ordered after and before `stream`; the remainder of
 * nsteps is launched directly. */
int fdtd_batch_run_graph(FdtdBatch* batch, int64_t q0, int64_t nsteps, int64_t steps_per_graph, void* stream);

/* Maximum bytes of batch tables that fdtd_batch_run_static can hold. */
#define FDTD_STATIC_BATCH_BYTES (4u << 20)

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (harnesses that forbid hipMalloc, tight serving loops): the same
 * work as fdtd_batch_create + fdtd_batch_set_launch + fdtd_batch_run over
 * the n finalized plans, and the same results, with the batch's tables
 * (per-grid arguments, work list, time index) kept in a device array of
 * FDTD_STATIC_BATCH_BYTES that the library holds on each GPU it runs on
 * until the process ends, instead of in memory it allocates. No device
 * memory is allocated or freed. The tables of a launch shape that need more
 * than FDTD_STATIC_BATCH_BYTES are refused (error -6); the 64-grid,
 * 132-million-cell batch needs about 2.2 MB at 256 threads and 1 plane.
 * The call copies the tables on `stream` and waits for that stream (work
 * queued on it before, then the copy), so it is not for graph capture.
 * Every call on a GPU shares that GPU's array and leaves its tables there:
 * calls that overlap on one GPU (other host threads, or other streams
 * whose work is still running) race. */
int fdtd_batch_run_static(int n, FdtdPlan* const* plans, int block_threads, int cx, int workgroups,
                          int64_t q0, int64_t nsteps, void* stream);

/* ---- Differentiation (reverse mode) -------------------------------------
 * The reference differentiates through its torch backend (fdtd/backend.py:
 * 207-361 with autograd enabled): torch records every array operation of
 * Grid.update_E/update_H and replays their transposes. These entry points
 * replace that for one objective, J = sum over the steps of sum over the
 * cells of an objective box of |E|^2 (E after each step), and its gradient
 * with respect to the inverse permittivity of a design box:
 *   fdtd_run_recorded  forward steps (as fdtd_run) that also record, after
 *                      each E half step, curl H inside the design box and E
 *                      inside the objective box;
 *   fdtd_adjoint_run   the transposed steps in reverse order over such a
 *                      record, accumulating dJ/d(inverse permittivity).
 * Periodic boundaries (registered after every PML face) and anisotropic
 * objects (not overlapping an absorbing one) are transposed too. Rejected:
 * FDTD_F32C and complex storage, a PML face registered after a periodic
 * boundary, more than one PML face per side, soft sources. */
typedef struct FdtdAdjoint FdtdAdjoint;

/* Elements of the adjoint CPML state buffer (plan dtype) that
 * fdtd_adjoint_create needs for this plan: 8 * (cells of each face slab). */
int fdtd_adjoint_psi_size(const FdtdPlan* plan, int64_t* elements);

/* Adjoint state over a finalized plan. Device arrays, plan dtype, borrowed:
 *   gcoef, dcoef  (3, Nx, Ny, Nz): the E update as E' = dcoef E + gcoef
 *                 curl_H(H) + CPML + sources, composed by the caller from the
 *                 grid and the objects (S inv_eps outside objects, S inv_obj
 *                 inside a plain object, absorbing objects scale both;
 *                 the 3x3 updates of anisotropic objects are not in gcoef:
 *                 the library adds their transposes from the plan's object
 *                 arrays); dcoef may be NULL (all ones)
 *   lamE, lamH    (3, Nx, Ny, Nz) adjoint fields, zeroed by the caller
 *   psi_adj       fdtd_adjoint_psi_size elements, zeroed by the caller
 *   grad          float64 (3, cells of the design box), z fastest, zeroed by
 *                 the caller; fdtd_adjoint_run adds to it
 * Boxes are half-open {x0, x1, y0, y1, z0, z1} inside the grid. */
int fdtd_adjoint_create(FdtdPlan* plan, const void* gcoef, const void* dcoef, void* lamE, void* lamH, void* psi_adj,
                        const int32_t design_box[6], const int32_t objective_box[6], double* grad,
                        FdtdAdjoint** out);

int fdtd_adjoint_destroy(FdtdAdjoint* adj);

/* nsteps forward steps from time index q0 through the adjoint's plan. After
 * the E half step of local step s: curl_records[s] = curl H inside the design
 * box (3, cells) and plane_records[s] = E inside the objective box (3, cells),
 * plan dtype. curl_records may be NULL (only the objective box is recorded).
 * The records are written by the step kernels themselves (no extra launch). */
int fdtd_run_recorded(FdtdAdjoint* adj, int64_t q0, int64_t nsteps, void* curl_records, void* plane_records,
                      void* stream);

/* The transposed steps for local steps nsteps-1 down to 0 of a record written
 * by fdtd_run_recorded: with lamE/lamH/psi_adj holding dJ/d(state after the
 * last of these steps), leave dJ/d(state before the first) and add the
 * segment's share to grad. Segments are processed latest first. lamE is kept
 * one factor behind: the dcoef of the last transposed E update is applied by
 * the next call, so with absorbing objects the final lamE lacks it. */
int fdtd_adjoint_run(FdtdAdjoint* adj, int64_t nsteps, const void* curl_records, const void* plane_records,
                     void* stream);

/* x planes per workgroup of the transposed kernels (0, the default: the
 * plan's choice for the main kernels). A tuning knob with no counterpart in
 * the reference: when the transposed steps run beside a recomputation on
 * another stream, fewer and longer workgroups leave room for it. */
int fdtd_adjoint_set_launch(FdtdAdjoint* adj, int cx);

/* Measurement hook, no counterpart in the reference (used by bench.py to
 * price the dominant kernels against the HBM roofline): launch the fused main
 * kernels of a step, E then H, `reps` times (at most 4096) on `stream` with a
 * HIP event recorded after each launch, wait for the last one and store the
 * average milliseconds of the E and of the H launches in *ms_E and *ms_H.
 * The launches advance the fields (without periodic stages, sources or
 * detectors), so use it on a scratch state. */
int fdtd_bench_main_kernel(FdtdPlan* plan, int64_t reps, void* stream, float* ms_E, float* ms_H);

/* Measurement hook, no counterpart in the reference: copy `bytes` (a multiple
 * of 16, both pointers 16-byte aligned) from src to dst `reps` times with a
 * plain streaming kernel between two HIP events on `stream`; *ms receives the
 * elapsed milliseconds. The practical HBM bandwidth next to the 8 TB/s peak. */
int fdtd_bench_copy(void* dst, const void* src, int64_t bytes, int64_t reps, void* stream, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* FDTD_HIP_H */
