/*
 * drn_oracle.h -- CPU restatement of DrNUM's per-patch explicit update, used
 * ONLY as test infrastructure (the checker for the HIP path, and the timed CPU
 * baseline).  Nothing in the product links, loads or calls this code.
 *
 * The same source (drn_oracle.c) is compiled twice: once with REAL=float
 * (symbols o32_*), matching the reference's default build (drnum.h:43), and
 * once with REAL=double (symbols o64_*), matching the FP64 branch of drnum.h.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root, src/drnumlib/ omitted where obvious).
 *
 * Pinning status of each part (what independent check fixes it) is listed in
 * the header comment of drn_oracle.c and in tests/test_oracle_*.py.
 */
#ifndef DRN_ORACLE_H
#define DRN_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* reconstruction (reconstruction/upwind1.h, upwind2.h) */
enum { O_RECON_UPWIND1 = 0, O_RECON_UPWIND2 = 1 };
/* limiters (reconstruction/vanalbada.h, minmod.h, secondorder.h) */
enum { O_LIM_VANALBADA = 0, O_LIM_MINMOD = 1, O_LIM_SECONDORDER = 2 };
/* wall functor at a patch face (the xWallM/xWallP/... surface):
 *  SLIP         fluxes/compressibleslipflux.h:34-98
 *  FARFIELD     fluxes/compressiblefarfieldflux.h:93-181 with Upwind1 (as JetFlux's farfield_t)
 *  TRANSMISSIVE not in the reference: Euler flux of the boundary cell state
 *  JET          applications/drnumJetDemo/jetflux.h:158-192 (top-hat inflow), generalised to any face
 *  NONE         face flux left at zero (used by faces whose cells are all overwritten by the exchange)  */
enum { O_BC_SLIP = 0, O_BC_FARFIELD = 1, O_BC_TRANSMISSIVE = 2, O_BC_JET = 3, O_BC_NONE = 4 };

/* parameters per face (16 doubles):
 *  FARFIELD: [0..4] = p, T, u, v, w of the far state
 *  JET:      [0] c_a, [1] c_b   jet centre in the two tangential patch-local coordinates
 *            [2] radius, [3] normal jet speed (positive into the domain), [4] T_jet,
 *            [5] normal far speed, [6] T_far, [7] noise amplitude (fraction of jet speed),
 *            [8] noise seed, [9] global index offset of tangential axis a, [10] of axis b,
 *            [11] p_jet, [12] supersonic flag (nonzero: p = p_jet inside the radius)  */
#define O_BC_NPAR 16

/* per-cell flags: Patch::deactivate (patch_common.h:505, the update skips the
 * cell, cartesianiterator.h:315) and m_IsInsideCell (split faces,
 * patch_common.h:500-503, upwind2.h:41) */
enum { O_CELL_INACTIVE = 1, O_CELL_INSIDE = 2 };

typedef struct o_patch_desc {
  int32_t ni, nj, nk;           /* cells per axis, seek layers included          */
  int32_t seek[6];              /* seek layers: imin, imax, jmin, jmax, kmin, kmax */
  double  origin[3];            /* patch origin (inertial frame)                 */
  double  length[3];            /* Lx, Ly, Lz                                     */
  int32_t bc[6];                /* face functor: xm, xp, ym, yp, zm, zp           */
  double  bc_par[6][O_BC_NPAR];
  double  base_i[3];            /* base vectors of the patch's i and j axes in the */
  double  base_j[3];            /* inertial frame (patch.cpp:63-85); (1,0,0), (0,1,0) axis aligned */
} o_patch_desc;

typedef struct o32_grid o32_grid;
typedef struct o64_grid o64_grid;

#define O_DECLARE(PFX, REAL, GRID)                                                              \
  /* perfectgas.h:56-64 */                                                                      \
  void PFX##prim_to_cons(REAL p, REAL T, REAL u, REAL v, REAL w, REAL* var);                     \
  /* reconstruction limiters */                                                                 \
  REAL PFX##limiter(int lim, REAL d1, REAL d2);                                                  \
  /* fluxes/ausmplus.h:39-106 on two reconstructed states, dir 0/1/2 = x/y/z */                 \
  void PFX##ausm_flux(const REAL* var_l, const REAL* var_r, int dir, REAL A, REAL* flux);        \
  /* perfectgas.h:100-111 Sutherland viscosity of a conserved state */                         \
  REAL PFX##gas_mu(const REAL* var);                                                            \
  /* fluxes/compressibleviscflux.h:33-268: face between var_l and var_r along dir; gl_a/gr_a   \
     and gl_b/gr_b are the cell gradients along the first and second tangential axis of the    \
     left and right cell, D = 1/delta along dir; subtracts from flux[1..4] */                  \
  void PFX##visc_face(const REAL* var_l, const REAL* var_r, const REAL* gl_a, const REAL* gr_a,  \
                      const REAL* gl_b, const REAL* gr_b, REAL D, int dir, REAL A, REAL* flux);  \
  /* grid object */                                                                             \
  GRID* PFX##grid_create(int n_patches, const o_patch_desc* pd, int recon, int lim);            \
  void  PFX##grid_destroy(GRID* g);                                                             \
  int   PFX##grid_num_patches(const GRID* g);                                                   \
  size_t PFX##grid_patch_cells(const GRID* g, int p);                                           \
  REAL* PFX##grid_field(GRID* g, int p, int field);                                             \
  REAL  PFX##grid_dx(const GRID* g, int p, int axis);                                           \
  void  PFX##grid_set_const(GRID* g, int field, const REAL* var);                               \
  /* iterators/cartesianiterator.h:87-325 for every patch: residual + update field0=field1+f*R */ \
  void  PFX##grid_compute(GRID* g, REAL factor);                                                \
  /* residual only of one patch (scratch of the iterator), 5*N values */                       \
  void  PFX##grid_residual(GRID* g, int p, REAL* res);                                          \
  void  PFX##grid_copy_field(GRID* g, int src, int dst);                                        \
  /* donor tables: patchgrid.cpp:119-259 (+ periodic images, not in the reference) */          \
  int   PFX##grid_compute_dependencies(GRID* g, const double* periodic);                        \
  int   PFX##grid_num_receivers(const GRID* g, int p);                                          \
  int   PFX##grid_num_no_donor(const GRID* g, int p);                                           \
  const int64_t* PFX##grid_receivers(const GRID* g, int p);                                     \
  int   PFX##grid_num_donors(const GRID* g, int p);                                             \
  /* donor d of patch p: out[0]=donor patch, [1]=num receiver cells, [2]=stride,               \
     [3]=receiver start in concat list, [4]=donor weight/index start */                         \
  void  PFX##grid_donor_info(const GRID* g, int p, int d, int64_t* out);                        \
  int64_t PFX##grid_num_concat(const GRID* g, int p, int which);                                \
  const int64_t* PFX##grid_rec_concat(const GRID* g, int p);                                    \
  const int64_t* PFX##grid_donor_index(const GRID* g, int p);                                   \
  const REAL* PFX##grid_donor_weight(const GRID* g, int p);                                     \
  /* patch.cpp:573-653 for every patch */                                                       \
  void  PFX##grid_exchange(GRID* g, int field);                                                 \
  /* per-cell O_CELL_* flags of patch p (N bytes; NULL clears) */                              \
  void  PFX##grid_set_cell_flags(GRID* g, int p, const unsigned char* flags);                   \
  /* donor d of patch p: the reference's donor_t matrix axx..azz (row-major), which turns    \
   * the donor's momentum into the receiver's frame (patch.cpp:178-186, 627-641) */           \
  void  PFX##grid_donor_matrix(const GRID* g, int p, int d, double* a9);                        \
  /* patch p's this->inertial transform as Patch::xyzoCell applies it (patch.cpp:91-98):       \
   * a12 = matrix (row-major) then shift */                                                     \
  void  PFX##grid_patch_transform(const GRID* g, int p, double* a12);                           \
  /* level set: levelsetobject.cpp:44-388 with spheres (spherelevelset.cpp, combilevelsetor.cpp) */ \
  int   PFX##grid_levelset(GRID* g, int n_spheres, const double* spheres, int n_inner,          \
                           int n_outer, double min_inner_rel);                                  \
  const REAL* PFX##grid_levelset_g(const GRID* g, int p);                                       \
  int   PFX##grid_ls_num_cells(const GRID* g, int p, int layer, int outer);                     \
  /* per layer cell: cell, then (as REAL) G, gx, gy, gz; 8 donors, 8 weights */               \
  void  PFX##grid_ls_cells(const GRID* g, int p, int layer, int outer, int64_t* cells,          \
                           int64_t* donors, REAL* gvals, REAL* weights);                        \
  /* cpu_levelsetobjectbc.h:58-90 with LSOBCCompressibleEulerOp (op 0, the default) or     */ \
  /* LSOBCCompressibleSWallOp (op 1, lsobccompressibleswallop.h:48-116)                    */ \
  void  PFX##grid_levelset_fill(GRID* g, int field, int abuse_field);                           \
  void  PFX##grid_set_levelset_op(GRID* g, int op);                                             \
  /* Patch::computeVariableDifference (patch.cpp:733-746), in REAL as the reference */         \
  void  PFX##grid_variable_difference(const GRID* g, int p, int f1, int v1, int f2, int v2,      \
                                      double* max_norm, double* l2_norm);                        \
  /* rungekutta.cpp:27-42 (+ level-set fill as post operation when set up) */                   \
  void  PFX##grid_rk_step(GRID* g, REAL dt, int n_alpha, const REAL* alpha);                    \
  void  PFX##grid_set_threads(GRID* g, int n);                                                  \
  /* EaFlux composition (drnumCylinder/main.h:127-163): the viscous flux is added to every     \
     interior face flux after the Euler flux; off by default */                                 \
  void  PFX##grid_set_viscous(GRID* g, int on);                                                 \
  /* FP64 sums over core cells: mass, x/y/z momentum, energy */                                 \
  void  PFX##grid_core_sums(const GRID* g, int field, double* out);

O_DECLARE(o32_, float, o32_grid)
O_DECLARE(o64_, double, o64_grid)

#ifdef __cplusplus
}
#endif
#endif
