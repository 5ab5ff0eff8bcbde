// dnd.cpp -- DrNUM restart files (.dnd), PatchGrid::writeData / readData
// (patchgrid.cpp:563-610) and Patch::writeData / readData (patch.cpp:776-790).
//
// The reference writes through a QDataStream, which is big-endian:
//   int32 precision (4 or 8), then `real time`, then for every patch in grid
//   order the fieldSize() values of one field, variable-major ([var][cell]),
//   each as a 4-byte or 8-byte IEEE number according to the precision word.
// Reading a file of the other precision converts value by value, as the
// stream does when its floating-point precision differs from `real`.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/drnum_host.h"

namespace {

void put_be(std::vector<unsigned char>& out, const void* p, int n) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  for (int i = n - 1; i >= 0; --i) out.push_back(b[i]);
}

bool get_be(FILE* f, void* p, int n) {
  unsigned char tmp[8];
  if (fread(tmp, 1, (size_t)n, f) != (size_t)n) return false;
  unsigned char* b = static_cast<unsigned char*>(p);
  for (int i = 0; i < n; ++i) b[i] = tmp[n - 1 - i];
  return true;
}

}  // namespace

extern "C" {

int drnh_dnd_write(const char* path, int precision, double time, int n_patches, const void* const* fields,
                   const int64_t* counts) {
  if (!path || (precision != 4 && precision != 8) || n_patches < 0 || (n_patches > 0 && (!fields || !counts)))
    return -1;
  FILE* f = fopen(path, "wb");
  if (!f) return -2;
  std::vector<unsigned char> buf;
  const int32_t prec = precision;
  put_be(buf, &prec, 4);
  if (precision == 4) {
    const float t = (float)time;
    put_be(buf, &t, 4);
  } else {
    put_be(buf, &time, 8);
  }
  bool ok = fwrite(buf.data(), 1, buf.size(), f) == buf.size();
  for (int p = 0; p < n_patches && ok; ++p) {
    buf.clear();
    buf.reserve((size_t)counts[p] * (size_t)precision);
    if (precision == 4) {
      const float* v = static_cast<const float*>(fields[p]);
      for (int64_t i = 0; i < counts[p]; ++i) put_be(buf, v + i, 4);
    } else {
      const double* v = static_cast<const double*>(fields[p]);
      for (int64_t i = 0; i < counts[p]; ++i) put_be(buf, v + i, 8);
    }
    ok = fwrite(buf.data(), 1, buf.size(), f) == buf.size();
  }
  if (fclose(f) != 0) ok = false;
  return ok ? 0 : -2;
}

int drnh_dnd_read(const char* path, int precision, double* time, int n_patches, void* const* fields,
                  const int64_t* counts, int* file_precision) {
  if (!path || (precision != 4 && precision != 8) || n_patches < 0 || (n_patches > 0 && (!fields || !counts)))
    return -1;
  std::string name(path);
  if (name.size() < 4 || name.compare(name.size() - 4, 4, ".dnd") != 0) name += ".dnd";
  FILE* f = fopen(name.c_str(), "rb");
  if (!f) return -2;
  int32_t prec = 0;
  bool ok = get_be(f, &prec, 4) && (prec == 4 || prec == 8);
  double t = 0;
  if (ok) {
    if (prec == 4) {
      float tf;
      ok = get_be(f, &tf, 4);
      t = tf;
    } else {
      ok = get_be(f, &t, 8);
    }
  }
  for (int p = 0; p < n_patches && ok; ++p) {
    for (int64_t i = 0; i < counts[p] && ok; ++i) {
      double v;
      if (prec == 4) {
        float vf;
        ok = get_be(f, &vf, 4);
        v = vf;
      } else {
        ok = get_be(f, &v, 8);
      }
      if (precision == 4)
        static_cast<float*>(fields[p])[i] = (float)v;
      else
        static_cast<double*>(fields[p])[i] = v;
    }
  }
  // the file must hold exactly the grid's fields
  if (ok) {
    unsigned char extra;
    if (fread(&extra, 1, 1, f) == 1) ok = false;
  }
  fclose(f);
  if (!ok) return -3;
  if (time) *time = t;
  if (file_precision) *file_precision = prec;
  return 0;
}

}  // extern "C"
