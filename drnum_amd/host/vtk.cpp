// vtk.cpp -- VTK output of PatchGrid::writeToVtk (patchgrid.cpp:673-695) and
// CartesianPatch::createVtkDataSet (cartesianpatch.cpp:1661-1730), written
// without the VTK library as VTK XML files:
//   <base>.vtm                 vtkMultiBlockDataSet, one block per patch
//   <base>/<base>_<p>.vts      vtkStructuredGrid of the patch's core cells
// A patch block holds, as the reference's data set does, the node points of the
// core cells (i*dx, j*dy, k*dz from the patch origin, float32, i fastest) and
// the post-processing quantities as float32 cell arrays (scalars, then
// 3-component vectors), in VTK's i-fastest cell order.  Data are raw appended
// little-endian blocks with a UInt64 byte count, as VTK's XML writers lay them
// out in "appended raw" mode.
#include <sys/stat.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/drnum_host.h"

namespace {

std::string base_name(const std::string& path) {
  const size_t s = path.find_last_of('/');
  return s == std::string::npos ? path : path.substr(s + 1);
}

}  // namespace

extern "C" {

int drnh_vtk_write_patch(const char* path, const int32_t* n3, const int32_t* seek6, const double* origin3,
                         const double* delta3, int n_scalars, const char* const* scalar_names,
                         const float* const* scalars, int n_vectors, const char* const* vector_names,
                         const float* const* vectors) {
  if (!path || !n3 || !seek6 || !origin3 || !delta3 || n_scalars < 0 || n_vectors < 0) return -1;
  if ((n_scalars > 0 && (!scalar_names || !scalars)) || (n_vectors > 0 && (!vector_names || !vectors))) return -1;
  const int ni = n3[0], nj = n3[1], nk = n3[2];
  const int i0 = seek6[0], i1 = ni - seek6[1], j0 = seek6[2], j1 = nj - seek6[3], k0 = seek6[4], k1 = nk - seek6[5];
  if (ni < 1 || nj < 1 || nk < 1 || i0 < 0 || j0 < 0 || k0 < 0 || i1 <= i0 || j1 <= j0 || k1 <= k0) return -1;
  const size_t ci = (size_t)(i1 - i0), cj = (size_t)(j1 - j0), ck = (size_t)(k1 - k0);
  const size_t ncell = ci * cj * ck, npts = (ci + 1) * (cj + 1) * (ck + 1);
  const size_t njk = (size_t)nj * (size_t)nk;

  // points: node (i, j, k) of the core at origin + (i, j, k) * delta, i fastest
  std::vector<float> pts(3 * npts);
  size_t q = 0;
  for (int k = k0; k <= k1; ++k)
    for (int j = j0; j <= j1; ++j)
      for (int i = i0; i <= i1; ++i) {
        pts[q++] = (float)(origin3[0] + i * delta3[0]);
        pts[q++] = (float)(origin3[1] + j * delta3[1]);
        pts[q++] = (float)(origin3[2] + k * delta3[2]);
      }
  // cell arrays in i-fastest order from the [cell = i*nj*nk + j*nk + k] layout
  std::vector<std::vector<float>> arrays;
  std::vector<std::string> names;
  std::vector<int> ncomp;
  for (int s = 0; s < n_scalars; ++s) {
    std::vector<float> a(ncell);
    size_t id = 0;
    for (int k = k0; k < k1; ++k)
      for (int j = j0; j < j1; ++j)
        for (int i = i0; i < i1; ++i) a[id++] = scalars[s][(size_t)i * njk + (size_t)j * nk + k];
    arrays.push_back(std::move(a));
    names.push_back(scalar_names[s]);
    ncomp.push_back(1);
  }
  for (int v = 0; v < n_vectors; ++v) {
    std::vector<float> a(3 * ncell);
    const size_t N = (size_t)ni * njk;
    size_t id = 0;
    for (int k = k0; k < k1; ++k)
      for (int j = j0; j < j1; ++j)
        for (int i = i0; i < i1; ++i) {
          const size_t c = (size_t)i * njk + (size_t)j * nk + k;
          for (int d = 0; d < 3; ++d) a[id++] = vectors[v][d * N + c];
        }
    arrays.push_back(std::move(a));
    names.push_back(vector_names[v]);
    ncomp.push_back(3);
  }

  FILE* f = fopen(path, "wb");
  if (!f) return -2;
  char ext[128];
  snprintf(ext, sizeof(ext), "0 %zu 0 %zu 0 %zu", ci, cj, ck);
  std::string x;
  x += "<?xml version=\"1.0\"?>\n";
  x += "<VTKFile type=\"StructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n";
  x += std::string("  <StructuredGrid WholeExtent=\"") + ext + "\">\n";
  x += std::string("    <Piece Extent=\"") + ext + "\">\n";
  x += "      <PointData>\n      </PointData>\n";
  x += "      <CellData>\n";
  uint64_t off = 0;
  for (size_t a = 0; a < arrays.size(); ++a) {
    x += "        <DataArray type=\"Float32\" Name=\"" + names[a] + "\"";
    if (ncomp[a] != 1) x += " NumberOfComponents=\"" + std::to_string(ncomp[a]) + "\"";
    x += " format=\"appended\" offset=\"" + std::to_string(off) + "\"/>\n";
    off += 8 + arrays[a].size() * sizeof(float);
  }
  x += "      </CellData>\n";
  x += "      <Points>\n";
  x += "        <DataArray type=\"Float32\" Name=\"Points\" NumberOfComponents=\"3\" format=\"appended\" offset=\"" +
       std::to_string(off) + "\"/>\n";
  x += "      </Points>\n";
  x += "    </Piece>\n  </StructuredGrid>\n";
  x += "  <AppendedData encoding=\"raw\">\n   _";
  bool ok = fwrite(x.data(), 1, x.size(), f) == x.size();
  auto block = [&](const std::vector<float>& a) {
    const uint64_t nb = a.size() * sizeof(float);
    ok = ok && fwrite(&nb, sizeof(nb), 1, f) == 1;
    ok = ok && (a.empty() || fwrite(a.data(), sizeof(float), a.size(), f) == a.size());
  };
  for (const auto& a : arrays) block(a);
  block(pts);
  const char* tail = "\n  </AppendedData>\n</VTKFile>\n";
  ok = ok && fwrite(tail, 1, strlen(tail), f) == strlen(tail);
  if (fclose(f) != 0) ok = false;
  return ok ? 0 : -2;
}

int drnh_vtk_write_multiblock(const char* base_path, int count, int n_patches, char* block_dir, int cap) {
  if (!base_path || n_patches < 0 || !block_dir || cap < 1) return -1;
  std::string base(base_path);
  if (count >= 0) {
    char num[16];
    snprintf(num, sizeof(num), "_%06d", count);
    base += num;
  }
  const std::string leaf = base_name(base);
  if (base.size() + 1 > (size_t)cap) return -2;
  // the blocks live in the directory <base>/ as <leaf>_<p>.vts
  if (mkdir(base.c_str(), 0755) != 0) {
    struct stat st;
    if (stat(base.c_str(), &st) != 0 || !S_ISDIR(st.st_mode)) return -2;
  }
  FILE* f = fopen((base + ".vtm").c_str(), "wb");
  if (!f) return -2;
  std::string x;
  x += "<?xml version=\"1.0\"?>\n";
  x += "<VTKFile type=\"vtkMultiBlockDataSet\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n";
  x += "  <vtkMultiBlockDataSet>\n";
  for (int p = 0; p < n_patches; ++p)
    x += "    <DataSet index=\"" + std::to_string(p) + "\" file=\"" + leaf + "/" + leaf + "_" + std::to_string(p) +
         ".vts\"/>\n";
  x += "  </vtkMultiBlockDataSet>\n</VTKFile>\n";
  bool ok = fwrite(x.data(), 1, x.size(), f) == x.size();
  if (fclose(f) != 0) ok = false;
  if (!ok) return -2;
  memcpy(block_dir, base.c_str(), base.size() + 1);
  return (int)base.size();
}

}  // extern "C"
