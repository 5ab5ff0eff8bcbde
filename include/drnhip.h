/*
 * drnhip.h -- C ABI of libdrnhip.so, the MI355X (gfx950) implementation of
 * DrNUM's per-patch explicit compressible update.
 *
 * The reference drives this path through C++ templates and CUDA kernels:
 *   GPU_PatchIterator / GPU_CartesianIterator (src/drnumlib/iterators/gpu_patchiterator.h:31-233,
 *   gpu_cartesianiterator.h:29-304), the flux-operator functors (fluxes/ausmplus.h:27-109 with
 *   reconstruction/upwind2.h + vanalbada.h / minmod.h, perfectgas.h), the donor exchange
 *   (Patch::accessDonorDataDirect, patch.cpp:573-653) and the level-set post operation
 *   (CPU_LevelSetObjectBC, cpu_levelsetobjectbc.h:58-90).
 * A C ABI cannot carry those templates, so the operator is described by enums
 * (the host maps each supported OP type to them) and patches by POD descriptors.
 *
 * Conventions: every int-returning call returns 0 on success or a negative
 * DRN_E* code; drn_last_error() gives the message.  The library never exits.
 * One context = one device + one HIP stream; calls are asynchronous on that
 * stream unless stated; drn_sync() is the barrier.  Calls on one context must
 * come from one host thread.  Host buffers are borrowed only for the duration
 * of an upload/download call.
 *
 * Data layout on the device: a context holds three state buffers (logical
 * fields 0, 1, 2 of the reference, patch_common.h:72-95).  Each buffer is
 * variable-major over all patches: variable v of patch p, cell l = i*NJ*NK +
 * j*NK + k, lives at buffer + v*stride + offset(p) + l, with offset(p) aligned
 * to 64 cells.  Host transfers use the reference's per-patch layout: 5 blocks
 * of N values, one per variable.
 */
#ifndef DRNHIP_H
#define DRNHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct drn_ctx drn_ctx;

/* state precision of a context */
enum { DRN_F32 = 0, DRN_F64 = 1 };
/* reconstruction: reconstruction/upwind1.h:29-39, upwind2.h:29-52 */
enum { DRN_RECON_UPWIND1 = 0, DRN_RECON_UPWIND2 = 1 };
/* limiter: reconstruction/vanalbada.h:28-33, minmod.h:28-31, secondorder.h:28-31 */
enum { DRN_LIM_VANALBADA = 0, DRN_LIM_MINMOD = 1, DRN_LIM_SECONDORDER = 2 };
/* face functor of a patch face (xWallM/xWallP/yWallM/yWallP/zWallM/zWallP):
 *   SLIP         fluxes/compressibleslipflux.h:34-98
 *   FARFIELD     fluxes/compressiblefarfieldflux.h:93-181 (Upwind1 projection, as JetFlux's farfield_t)
 *   TRANSMISSIVE zero-gradient outflow (no counterpart in the reference)
 *   JET          applications/drnumJetDemo/jetflux.h:158-192, top-hat inflow with seeded velocity noise
 *   NONE         no wall flux (faces whose cells are all exchange receivers) */
enum { DRN_BC_SLIP = 0, DRN_BC_FARFIELD = 1, DRN_BC_TRANSMISSIVE = 2, DRN_BC_JET = 3, DRN_BC_NONE = 4 };
#define DRN_BC_NPAR 16

/* launch forms of a stage */
enum {
  DRN_LAUNCH_BATCHED = 0,   /* one launch over the tile work-list of all patches */
  DRN_LAUNCH_PER_PATCH = 1  /* one launch per patch (the plain drop-in form)    */
};

/* error codes */
enum {
  DRN_OK = 0,
  DRN_E_ARG = -1,      /* invalid argument (sizes, ids, enums)   */
  DRN_E_HIP = -2,      /* HIP runtime error                       */
  DRN_E_STATE = -3,    /* call out of order (e.g. before layout)  */
  DRN_E_NOMEM = -4,    /* device allocation failed                */
  DRN_E_RANGE = -5     /* table entry outside the patch or grid   */
};

typedef struct drn_patch_desc {
  int32_t ni, nj, nk;            /* cells per axis (seek layers included)                   */
  int32_t seek[6];               /* seek layers imin, imax, jmin, jmax, kmin, kmax (core)    */
  double dx, dy, dz;             /* spacing as the host computed it (L/N in the state type)  */
  int32_t bc[6];                 /* DRN_BC_* per face xm, xp, ym, yp, zm, zp                 */
  double bc_par[6][DRN_BC_NPAR]; /* FARFIELD: [0..4] far state in CONSERVED variables;       *
                                  * JET: [0] c_a [1] c_b [2] radius [3] jet speed [4] T_jet     *
                                  *      [5] far normal speed [6] T_far [7] noise amplitude    *
                                  *      [8] seed [9],[10] global index offsets (axes a, b)    *
                                  *      [11] p_jet [12] supersonic: nonzero takes p = p_jet    *
                                  *      inside the radius (jetflux.h:176-178), else p of the   *
                                  *      boundary cell                                          */
} drn_patch_desc;

/* ---- context ------------------------------------------------------------ */
/* Replaces the device selection and stream of GPU_PatchIterator(OP, cuda_device, thread_limit)
 * (gpu_patchiterator.h:68-83). */
int drn_ctx_create(int device, int precision, drn_ctx** out);
void drn_ctx_destroy(drn_ctx* ctx);
const char* drn_last_error(const drn_ctx* ctx);
int drn_precision(const drn_ctx* ctx);

/* The operator: AusmPlus<5, Upwind1|Upwind2<5, LIM>, PerfectGas> (ausmplus.h:27-109). */
int drn_set_operator(drn_ctx* ctx, int recon, int limiter);
/* on = 1 adds CompressibleViscFlux<5, PerfectGas> (compressibleviscflux.h:27-268)
 * to every interior face flux after the Euler flux, as the applications' flux
 * classes compose them (EaFlux::xField, drnumCylinder/main.h:127-163; JetFlux,
 * drnumJetDemo/jetflux.h:100-139); the face functors are unchanged.  Every
 * tiling choice then runs the z-shared cell kernel (27 or 28 by list order).
 * Default 0. */
int drn_set_viscous(drn_ctx* ctx, int on);

/* ---- patches ------------------------------------------------------------ */
/* Replaces PatchIterator::addPatch (patchiterator.h:124-128) + the lazy device
 * mirror GPU_Patch (gpu_patch.h:43-129).  Patches are added before drn_layout. */
int drn_patch_add(drn_ctx* ctx, const drn_patch_desc* desc, int* id);
/* Allocate the three state buffers for all added patches (gpu_patch.h:57-128).
 * One variable array of the grid (all patches, each padded to 64 cells) must
 * stay below 2^32 bytes: 2^30 stored cells in FP32, 2^29 in FP64. */
int drn_layout(drn_ctx* ctx);
int drn_num_patches(const drn_ctx* ctx);
size_t drn_patch_cells(const drn_ctx* ctx, int id);
/* Host <-> device copy of one field of one patch (GPU_Patch::copyToDevice /
 * copyFromDevice, gpu_patch.h:131-140; updateDevice/updateHost, gpu_patchiterator.h:85-126).
 * `host` holds 5*N values of the context precision, variable-major. Synchronous. */
int drn_patch_upload(drn_ctx* ctx, int id, int field, const void* host);
int drn_patch_download(drn_ctx* ctx, int id, int field, void* host);
/* PatchIterator::copyField (patchiterator.h:96-101) for all patches. */
int drn_copy_field(drn_ctx* ctx, int src, int dst);
/* Exchange the device buffers behind two logical fields (no data movement).
 * The host iterator uses it to give drn_stage the in-place semantics of
 * CartesianIterator::compute (field 0 = field 1 + factor/dV * Res(field 0)). */
int drn_swap_fields(drn_ctx* ctx, int a, int b);

/* Per-cell flags of one patch: `flags` holds one byte per cell in the patch's
 * cell order, NULL (or all zero) clears them.
 *   DRN_CELL_INACTIVE  Patch::deactivate (patch_common.h:505): the stage leaves
 *                      the cell's state as it is (CartesianIterator::compute,
 *                      cartesianiterator.h:315, updates active cells only);
 *   DRN_CELL_INSIDE    m_IsInsideCell (patch_common.h:56): a face between an
 *                      inside and an outside cell is a split face
 *                      (Patch::isSplitFace, patch_common.h:500-503), and the
 *                      Upwind2 reconstruction of a cell whose face toward its
 *                      outer neighbour i0 is a split face stays first order
 *                      (upwind2.h:41).
 * While any patch has flags every tiling choice runs the z-shared cell kernel
 * (27 or 28 by list order) with the flag tests.  After drn_layout; drops a
 * captured graph. */
enum { DRN_CELL_INACTIVE = 1, DRN_CELL_INSIDE = 2 };
int drn_set_cell_flags(drn_ctx* ctx, int patch, const uint8_t* flags);

/* ---- one RK stage: flux + residual + update ------------------------------ */
/* GPU_CartesianIterator::compute (gpu_cartesianiterator.h:219-304) and
 * CartesianIterator::compute (cartesianiterator.h:87-325): for every patch
 *   field dst = field base + factor/dV * Res(field src)
 * where Res sums the AUSM+ face fluxes of all cell faces and the face functors
 * at the patch faces.  dst must differ from src; dst may equal base.
 * `launch` is DRN_LAUNCH_BATCHED or DRN_LAUNCH_PER_PATCH. */
int drn_stage(drn_ctx* ctx, int src, int base, int dst, double factor, int launch);

/* ---- donor exchange ------------------------------------------------------ */
/* Tables of one receiver patch in the reference's direct transfer-list form
 * (Patch::buildDonorTransferData, patch.cpp:128-218; donor_t, donor_t.h:28-45):
 *   n_donors donors; donor d is patch donor_patch[d] and serves
 *   donor_info[4d+0] receiver cells with stride donor_info[4d+1], its receivers
 *   start at donor_info[4d+2] in rec_concat and its (index, weight) pairs at
 *   donor_info[4d+3] in donor_index / donor_weight.  rec_unique lists every
 *   receiver cell once (zeroed before the donor sums).  Cell indices are local
 *   to their patch.  Weights are FP64 and rounded to the state precision. */
int drn_exchange_set(drn_ctx* ctx, int recv_id, int n_donors, const int32_t* donor_patch,
                     const int64_t* donor_info, int64_t n_unique, const int64_t* rec_unique,
                     int64_t n_rec_concat, const int64_t* rec_concat, int64_t n_wi,
                     const int64_t* donor_index, const double* donor_weight);
/* The turn of vector variables of each donor of receiver patch recv_id, set
 * after drn_exchange_set for that patch (which clears it): turn[9d..9d+8] is
 * donor d's matrix axx, axy, axz, ayx, ..., azz of donor_t (donor_t.h:36-44,
 * Patch::buildDonorTransferData, patch.cpp:178-186), rounded to the state
 * precision.  The exchange turns the momentum (variables 1-3, the vector
 * variable PatchGrid::defineVectorVar(1) declares, drnumJetDemo/main.h:99) of
 * each donor's weighted sum with it before adding the sum to the receiver
 * (Patch::accessDonorDataDirect, patch.cpp:617-645).  A matrix with exact 1 on
 * the diagonal and zeros elsewhere (aligned patches) changes no finite value
 * and is skipped.  Not calling it leaves every donor of the patch aligned. */
int drn_exchange_set_turn(drn_ctx* ctx, int recv_id, int n_donors, const double* turn);
/* Upload all tables set so far into the device gather lists.  Fails with
 * DRN_E_RANGE when a donor cell is also a receiver cell: the reference fills
 * receivers patch after patch, the device all at once, so such tables would
 * give order-dependent results (the reference's donor search never produces
 * them: donors lie in the core of their patch). */
int drn_exchange_commit(drn_ctx* ctx);
/* PatchIterator::copyDonorData(field) (patchiterator.h:103-108, gpu_patchiterator.h:191-233). */
int drn_exchange(drn_ctx* ctx, int field);
/* No counterpart in the reference, whose copyDonorData copies every receiver
 * after the stage.  With fusion on (the default), drn_rk_step and
 * drn_rk_graph_run let the cell stage kernels (the tiling kinds of choices 11,
 * 20, 27, 28, 33, 34, 35) fill part of the exchange themselves: the z-face seek
 * layers of a patch's x/y core columns, one side at a time, where
 * drn_exchange_commit found that side to be an aligned unit-weight copy of
 * one donor patch.  A donor thread writes 0 + its new value into its receiver
 * (the exchange's own sum), the receiver thread skips its store, and the
 * exchange after the stage runs on the other receivers.  The results are
 * bit-identical to fusion off; drn_stage and drn_exchange never fuse.
 * Switching drops a captured graph. */
int drn_set_exchange_fusion(drn_ctx* ctx, int on);
/* Receivers that the stage kernels of drn_rk_step fill themselves for this
 * launch form (DRN_LAUNCH_*) under the current tiling and fusion switch: 0 when
 * fusion is off, the tiling's kernel does not fuse or no side qualified at
 * drn_exchange_commit. */
int drn_exchange_fused_cells(const drn_ctx* ctx, int launch, int64_t* n);

/* ---- level-set boundary fill -------------------------------------------- */
/* Inner layer cells of one patch in the flat order of
 * LevelSetObjectBC::transferCellLayerData (levelsetobjectbc.cpp:38-134):
 * cell, unit normal (gx, gy, gz), 8 mirror donors and weights (levelsetobject.cpp:294-359). */
int drn_levelset_set(drn_ctx* ctx, int patch, int64_t n, const int64_t* cells, const double* normals,
                     const int64_t* donors, const double* weights);
int drn_levelset_commit(drn_ctx* ctx);
/* CPU_LevelSetObjectBC<OP>::operator() (cpu_levelsetobjectbc.h:58-90) on `field`,
 * with OP LSOBCCompressibleEulerOp (lsobccompressibleeulerop.h:49-144, the slip
 * wall) or LSOBCCompressibleSWallOp (lsobccompressibleswallop.h:48-116, the
 * simple wall) as drn_set_levelset_op chose; the mirror gather goes to device
 * scratch (the reference's abuse field). relax is 0.5 in the reference. */
int drn_levelset_fill(drn_ctx* ctx, int field, double relax);
/* The OP template argument of CPU_LevelSetObjectBC<DIM, OP> for drn_levelset_fill
 * and the fill inside drn_rk_step.  Default DRN_LS_SLIP.  Drops a captured graph. */
enum { DRN_LS_SLIP = 0, DRN_LS_SWALL = 1 };
int drn_set_levelset_op(drn_ctx* ctx, int op);

/* ---- time step ------------------------------------------------------------ */
/* RungeKutta::operator()(dt) (rungekutta.cpp:27-42): one low-storage RK step
 * with coefficients alpha[0..n-1], the exchange after every stage and the
 * level-set fill (if tables are set) as the post operation.  The state is in
 * field 0 before and after; fields 1 and 2 are used as stage storage.  The
 * exchange before a stage is repeated only when a post operation exists or
 * when the state was written since its last exchange (an upload, a copy, a
 * drn_stage or a fill into it); otherwise it would reproduce the same values. */
int drn_rk_step(drn_ctx* ctx, double dt, int n_alpha, const double* alpha, int launch);
/* `n_steps` drn_rk_step calls (batched launch form) replayed from HIP graphs.
 * A graph holds the steps until the field map returns to its start (two for an
 * odd stage count) and the physical buffers of the map it was captured from,
 * so one is captured per starting map and per (dt, alpha); a remainder of
 * fewer steps runs eagerly.  Results are bit-identical to drn_rk_step. */
int drn_rk_graph_run(drn_ctx* ctx, double dt, int n_alpha, const double* alpha, int n_steps);

/* ---- reductions (FP64 accumulators) -------------------------------------- */
/* out[0..4]: sum over core cells of (rho, rhou, rhov, rhow, rhoE) * dV. Synchronous. */
int drn_core_sums(drn_ctx* ctx, int field, double* out);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): drn_core_sums with the result written to d_out, 5 doubles of
 * DEVICE memory that the caller owns, asynchronously on the context stream
 * (drn_sync before reading it).  The per-chunk partial sums go to a fixed
 * device array of the library (40 MiB, held on every device the library runs
 * on until the process ends) instead of context memory, so the call allocates
 * nothing.  The sums equal drn_core_sums bit for bit.  At most 2^20 plane
 * chunks of 256 cells (about 2.7e8 stored cells): DRN_E_RANGE above.  Every
 * call on a device shares that array: calls that overlap on one device (two
 * contexts on two host threads, or two streams) race. */
int drn_core_sums_noalloc(drn_ctx* ctx, int field, double* d_out);
/* The diagnostics loop of the jet demo (applications/drnumJetDemo/main.h:188-210)
 * over all cells: out[0] = CFL_max = max over cells of max(|u|, |u+a|, |u-a|)
 * dt/dx with the x velocity and spacing only, evaluated in the state precision
 * as the reference does; out[1] = rho min, out[2] = rho max, out[3] = number of
 * non-finite values.  Synchronous. */
int drn_diagnostics(drn_ctx* ctx, int field, double dt, double* out);
/* No counterpart in the reference: the CFL number over all three axes,
 * max over cells and axes d of (|u_d| + a) dt/delta_d, evaluated in FP64, for
 * flows that are not aligned with x.  Synchronous. */
int drn_cfl_axes(drn_ctx* ctx, int field, double dt, double* cfl);
/* Patch::computeVariableDifference(field1, var1, field2, var2, max_norm, l2_norm)
 * (patch.cpp:733-746) over all cells of `patch`: max_norm = sqrt(max (b - a)^2),
 * l2_norm = sqrt(sum (b - a)^2), each square in the state type.  max_norm is the
 * reference's value bit for bit; the sum is accumulated in FP64 (the reference
 * accumulates in the state type).  Synchronous. */
int drn_variable_difference(drn_ctx* ctx, int patch, int field1, int var1, int field2, int var2, double* max_norm,
                            double* l2_norm);

/* ---- post-processing ----------------------------------------------------- */
/* CompressibleVariables<PerfectGas>::getScalar / getVector (compressiblevariables.h:81-105)
 * of every cell of one patch, evaluated on the device.  Bit q of `mask` selects
 * q = 0 Ma, 1 p, 2 T, 3 rho, 4 S = cp ln(T/ref_T) - R ln(p/ref_p), 5 u, 6 v, 7 w
 * (the reference defaults are ref_p = 1e5, ref_T = 300).  host_out receives the
 * selected quantities in increasing q, N values each, in the context
 * precision.  Only the selected quantities cross PCIe.  Synchronous. */
int drn_post_variables(drn_ctx* ctx, int field, int patch, int mask, double ref_p, double ref_T, void* host_out);

/* ---- tuning, timing, sync ------------------------------------------------ */
/* Stage-kernel work-list choices, per launch form (drn_num_tilings() of them):
 * 0-10  the marching kernel (planes staged in LDS; in fast arithmetic the
 *       slope form) with an i-range of the whole patch / 32 / 16 / 8 / 4 / 2
 *       planes per work item, in natural (patch, plane range, chunk) order or
 *       segmented per XCD (workgroup b runs on XCD b % 8, so each XCD gets one
 *       contiguous run of neighbouring chunks and shares their halos in L2);
 * 11    the cell-per-thread kernel;
 * 12-18 the column kernel (fast MUSCL van Albada / minmod only, otherwise the
 *       marching kernel): no LDS, z faces shared between lanes, with the
 *       same i-range / order variants;
 * 19    the cell-per-thread kernel with per-cell limited slopes (fast MUSCL
 *       van Albada / minmod only, otherwise 11);
 * 20    as 11, work list in 8 contiguous segments, one per XCD;
 * 21    as 19, work list in 8 contiguous segments, one per XCD;
 * 22-26 the wave-tile kernel: each wavefront owns an 8x8 (j,k) lane tile and
 *       marches over an i-range of the whole patch / 16 / 8 / 4 planes (XCD
 *       segmented) or 8 planes (natural order); every face is evaluated once
 *       per tile, y/z neighbours come from a wave-private LDS plane tile and
 *       lane shuffles, with no block-level barrier;
 * 27, 28 the cell-per-thread kernel with 63 cells per wavefront whose z faces
 *       are evaluated once and handed to the neighbouring lane (XCD segmented /
 *       natural order);
 * 29, 30 as 27 / 28 with 8 wavefronts per workgroup sharing their y faces
 *       through LDS (one barrier);
 * 31, 32 as 27 / 28 with two cells per thread along i, the x face between
 *       them evaluated once;
 * 33, 34 the work list of 27 / 28 in one persistent launch: as many
 *       workgroups as the device holds at once (a multiple of 8), each
 *       walking the list with a stride of the grid size;
 * 35    the kernel of 27 on a list that walks the patches joined by fused
 *       z-face pushes (drn_set_exchange_fusion) plane by plane together, then
 *       XCD segmented: a donor's pushed rows and the receiver's own rows meet
 *       in one L2.  Without fused pushes it is 27; one launch per patch runs 27.
 * 36-39 the z-shared cell kernel marching its 252 (j,k) cells over an i-range
 *       (whole patch / 32 / 16 / 8 planes per item, XCD segmented): each x
 *       face is evaluated once and its left state carried to the next plane,
 *       so a cell evaluates 4 faces (the march needs 3 or more cells per axis;
 *       thinner patches run 11);
 * 40, 41 as 36 / 38 capped at 8 waves per SIMD (FP32 fast arithmetic, else 36 / 38);
 * 42, 43 as 36 / 38 capped at 7 waves per SIMD (FP32 fast arithmetic, else 36 / 38);
 * 44-46 as 36 / 38 / 39 with the carried x slope and x face flux held in LDS
 *       instead of registers.
 * 47, 48 the z-shared cell kernel over 4 consecutive planes per work item
 *       (one wavefront of 63 cells per plane, the same (j,k) cells): a
 *       wavefront above the first plane takes its x face i-1/2 from the
 *       wavefront below through LDS (one barrier), 4.27 instead of 5.02 face
 *       evaluations per cell; on the list of 35 (push chains) / XCD segmented;
 * 49, 50 the same over 2 planes per work item (two wavefronts per plane), 4.52
 *       face evaluations per cell.
 * Default for both forms: 11; drn_autotune picks per device. */
int drn_num_tilings(void);
int drn_set_tiling(drn_ctx* ctx, int launch, int choice);
int drn_get_tiling(const drn_ctx* ctx, int launch);
/* Time every choice on one stage and the exchange after it, for the launch
 * form, on the current state (the state is not modified), and keep the
 * fastest; writes it to *choice if not NULL. */
int drn_autotune(drn_ctx* ctx, int launch, int* choice);
/* Arithmetic of the flux kernels: IEEE division and square root, or the fast
 * form (hardware reciprocal, single-division van Albada, exact 0/1 for the
 * supersonic pressure split).  Default: FAST for DRN_F32, IEEE for DRN_F64. */
enum { DRN_ARITH_IEEE = 0, DRN_ARITH_FAST = 1 };
int drn_set_arith(drn_ctx* ctx, int mode);
/* Average device time (ms) of the launches of one kernel class recorded since
 * the previous call, measured with HIP events on the context stream while
 * timing is enabled: which = 0 stage, 1 exchange, 2 level-set fill, 3 field copy. */
int drn_timing(drn_ctx* ctx, int which, double* avg_ms, int64_t* count);
int drn_timing_enable(drn_ctx* ctx, int on);
int drn_sync(drn_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
