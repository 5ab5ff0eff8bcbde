/*
 * drnum_host.h -- C ABI of libdrnum.so, the host side of the drop-in: patch
 * grid, grid-file reader, donor tables and level-set tables (FP64 geometry),
 * and the hand-over of both table sets to a libdrnhip context.
 *
 * The reference builds these on the host in C++ (PatchGrid::readGrid,
 * patchgrid.cpp:479-561; PatchGrid::computeDependencies, patchgrid.cpp:119-259;
 * LevelSetObject::update/extractBCellLayers, levelsetobject.cpp:44-388) and its
 * GPU iterator copies them to the device (GPU_Patch, gpu_patch.h:43-173).  This
 * C ABI is what a non-C++ caller (the ctypes tests and benchmark) binds; C++
 * callers use drnum_amd/host/grid.h and iterators.h directly.
 *
 * Return codes: 0 success, negative on error (message via drnh_last_error).
 */
#ifndef DRNUM_HOST_H
#define DRNUM_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "drnhip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct drnh_grid drnh_grid;

/* geometry precision: 1 = FP64 (product), 0 = FP32 (the reference's default `real`) */
drnh_grid* drnh_grid_create(int fp64_geometry);
void drnh_grid_destroy(drnh_grid* g);
const char* drnh_last_error(const drnh_grid* g);

/* CartesianPatch set-up (cartesianpatch.cpp:98-161): cells per axis including
 * seek layers, seek layers per face, origin and lengths; returns the patch id. */
int drnh_grid_add_patch(drnh_grid* g, int ni, int nj, int nk, const int32_t* seek6, const double* origin3,
                        const double* length3);
/* The base vectors of patch p's local i and j axes in the inertial frame
 * (Patch::readFromFile, patch.cpp:63-85, setupTransformation :265-278); a
 * patch added with drnh_grid_add_patch is axis aligned, (1,0,0) / (0,1,0).
 * Fails when a vector is zero or the two are parallel. */
int drnh_grid_set_base(drnh_grid* g, int p, const double* base_i3, const double* base_j3);
/* out6 = base_i, base_j of patch p as stored (geometry precision, widened) */
int drnh_grid_patch_base(const drnh_grid* g, int p, double* out6);
/* out12 = patch p's this -> inertial transform as Patch::xyzoCell applies it
 * (patch.cpp:91-98): the row-major matrix, then the shift */
int drnh_grid_patch_transform(const drnh_grid* g, int p, double* out12);
/* PatchGrid::readGrid (patchgrid.cpp:479-561). */
int drnh_grid_read(drnh_grid* g, const char* path, double scale);
int drnh_grid_num_patches(const drnh_grid* g);
/* out_i[9] = ni, nj, nk, seek[6]; out_d[9] = origin[3], length[3], dx, dy, dz */
int drnh_grid_patch_info(const drnh_grid* g, int p, int32_t* out_i, double* out_d);
/* face functors of a patch (DRN_BC_* and DRN_BC_NPAR parameters per face).
 * A farfield face takes the primitive state p, T, u, v, w in par[0..4], as
 * CompressibleFarfieldFlux::setFarfield does (compressiblefarfieldflux.h:74-77);
 * drnh_grid_to_device converts it to conserved variables in the state precision. */
int drnh_grid_set_faces(drnh_grid* g, int p, const int32_t* bc6, const double* par6x16);

/* ---- solver code strings and patch groups -------------------------------
 * CodeString (codestring.cpp:33-212) and PatchGroups (patchgroups.cpp:27-112),
 * as the reference's IteratorFeeder (iteratorfeeder.cpp:29-62) uses them to
 * route patches to iterators.  Strings are written NUL-terminated into
 * out[cap]; the functions return the string length, or -2 when cap is too
 * small, -1 on a bad argument. */
/* CodeString(text): the normalised code string */
int drnh_codestring(const char* text, char* out, int cap);
/* CodeString(a).compareCodeString(CodeString(b), ...): flags3 = allsame,
 * overruled, underruled; out = the word-wise merge ("" when they differ) */
int drnh_codestring_compare(const char* a, const char* b, int32_t* flags3, char* out, int cap);
/* code string of patch p as read from the grid file */
int drnh_grid_patch_codes(const drnh_grid* g, int p, char* out, int cap);
/* the grid's patch groups (filled by drnh_grid_read, exceptions allowed) */
int drnh_grid_num_patch_groups(const drnh_grid* g);
/* info3 = patch type, number of patches, solver exceptions (0/1); patches gets
 * the grid indices of the group's patches (info3[1] of them, may be NULL);
 * out = the group's code string */
int drnh_grid_patch_group(const drnh_grid* g, int group, int32_t* info3, int32_t* patches, char* out, int cap);

/* PatchGrid::computeDependencies; period[3] > 0 adds periodic images per axis
 * (not in the reference). */
int drnh_grid_compute_dependencies(drnh_grid* g, const double* period3, int threads);
int64_t drnh_grid_num_receivers(const drnh_grid* g, int p);
int64_t drnh_grid_num_no_donor(const drnh_grid* g, int p);
int drnh_grid_receivers(const drnh_grid* g, int p, int64_t* out);
int drnh_grid_num_donors(const drnh_grid* g, int p);
/* out[5] = donor patch, receiver cells, stride, receiver start, index/weight start */
int drnh_grid_donor_info(const drnh_grid* g, int p, int d, int64_t* out);
/* which = 0: concatenated receiver list length, 1: index/weight list length */
int64_t drnh_grid_num_concat(const drnh_grid* g, int p, int which);
int drnh_grid_rec_concat(const drnh_grid* g, int p, int64_t* out);
int drnh_grid_donor_index(const drnh_grid* g, int p, int64_t* out);
int drnh_grid_donor_weight(const drnh_grid* g, int p, double* out);
/* out9 = donor d's turn of the momentum into receiver p's frame, donor_t
 * axx..azz row-major (patch.cpp:178-186); handed to the device with
 * drn_exchange_set_turn by drnh_grid_upload_exchange */
int drnh_grid_donor_turn(const drnh_grid* g, int p, int d, double* out9);

/* Level set min_s(|x - c_s| - r_s) of spheres (xc, yc, zc, r) x n (spherelevelset.cpp,
 * combilevelsetor.cpp); layers as LevelSetObject(.., n_inner, n_outer, min_inner_rel, -1). */
int drnh_grid_levelset_spheres(drnh_grid* g, int n_spheres, const double* spheres, int n_inner, int n_outer,
                               double min_inner_rel);
int drnh_grid_levelset_values(const drnh_grid* g, int p, double* out);
int64_t drnh_grid_ls_num_cells(const drnh_grid* g, int p, int layer, int outer);
/* per cell: cell; gvals[4] = G, gx, gy, gz; donors[8]; weights[8] */
int drnh_grid_ls_cells(const drnh_grid* g, int p, int layer, int outer, int64_t* cells, double* gvals,
                       int64_t* donors, double* weights);

/* Restart files (.dnd): PatchGrid::writeData / readData (patchgrid.cpp:563-610,
 * patch.cpp:776-790).  Big-endian int32 precision (4 | 8), the time, then for
 * each patch in grid order counts[p] values of one field ([var][cell]).
 * `precision` is that of the caller's buffers (fields[p] are float* for 4,
 * double* for 8); a file of the other precision is converted on reading.
 * drnh_dnd_read appends ".dnd" when the name lacks it, and fails (-3) on a
 * short or over-long file.  0 on success, -1 bad argument, -2 I/O error. */
int drnh_dnd_write(const char* path, int precision, double time, int n_patches, const void* const* fields,
                   const int64_t* counts);
int drnh_dnd_read(const char* path, int precision, double* time, int n_patches, void* const* fields,
                  const int64_t* counts, int* file_precision);

/* PatchGrid::computeMinChLength (patchgrid.cpp:659-671) */
double drnh_grid_min_ch_length(const drnh_grid* g);

/* ---- VTK output -----------------------------------------------------------
 * PatchGrid::writeToVtk (patchgrid.cpp:673-695) with CartesianPatch::
 * createVtkDataSet (cartesianpatch.cpp:1661-1730), as VTK XML files written
 * without the VTK library: <base>[_<count, 6 digits>].vtm (a vtkMultiBlockDataSet
 * of n_patches blocks) and the block files <base>/<leaf>_<p>.vts. */
/* Writes the .vtm index (count < 0: no count suffix), creates the block
 * directory and returns its path in block_dir (the string length, -2 on I/O
 * error or a too small cap). */
int drnh_vtk_write_multiblock(const char* base_path, int count, int n_patches, char* block_dir, int cap);
/* One patch as a vtkStructuredGrid of its core cells (seek layers left out):
 * node points origin + (i, j, k) * delta (float32, i fastest), then the cell
 * arrays (float32, i fastest): n_scalars arrays of ni*nj*nk values and
 * n_vectors arrays of 3 consecutive ni*nj*nk blocks (x, y, z components), all
 * in the patch layout cell = i*nj*nk + j*nk + k. */
int drnh_vtk_write_patch(const char* path, const int32_t* n3, const int32_t* seek6, const double* origin3,
                         const double* delta3, int n_scalars, const char* const* scalar_names,
                         const float* const* scalars, int n_vectors, const char* const* vector_names,
                         const float* const* vectors);

/* Add every patch of the grid to a libdrnhip context (in grid order, with the
 * face functors set by drnh_grid_set_faces) and lay the context out. */
int drnh_grid_to_device(drnh_grid* g, drn_ctx* ctx);
/* Hand the donor tables (after compute_dependencies) to the context and commit. */
int drnh_grid_upload_exchange(drnh_grid* g, drn_ctx* ctx);
/* Hand the inner level-set layers (flat order patch, layer, list) to the context and commit. */
int drnh_grid_upload_levelset(drnh_grid* g, drn_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
