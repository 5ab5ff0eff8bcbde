// capi.cpp -- C ABI of the host library (include/drnum_host.h).
#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>
#include <string>

#include "../../include/drnum_host.h"
#include "grid.h"

using namespace drnum;

struct drnh_grid {
  bool fp64 = true;
  std::unique_ptr<PatchGrid<double>> gd;
  std::unique_ptr<PatchGrid<float>> gf;
  std::string err;
};

namespace {

template <typename F>
auto with(drnh_grid* g, F&& f) {
  return g->fp64 ? f(*g->gd) : f(*g->gf);
}
template <typename F>
auto with(const drnh_grid* g, F&& f) {
  return g->fp64 ? f(static_cast<const PatchGrid<double>&>(*g->gd)) : f(static_cast<const PatchGrid<float>&>(*g->gf));
}

bool bad_patch(const drnh_grid* g, int p) { return !g || p < 0 || p >= drnh_grid_num_patches(g); }

// PerfectGas::primitiveToConservative (perfectgas.h:56-64) in the state
// precision T; 0.5 is a double literal there, so the energy term is formed in
// double and rounded once.
template <typename T>
void prim_to_cons(T p, T temp, T u, T v, T w, T* var) {
  var[0] = p / (T(287.2) * temp);
  var[1] = var[0] * u;
  var[2] = var[0] * v;
  var[3] = var[0] * w;
  var[4] = T(var[0] * (T(717.5) * temp + 0.5 * (u * u + v * v + w * w)));
}

// The farfield functor keeps its state in conserved variables
// (CompressibleFarfieldFlux::setFarfield, compressiblefarfieldflux.h:74-77);
// the host keeps the primitive p, T, u, v, w the caller gave.
template <typename T>
void farfield_to_conserved(const double* prim, double* out) {
  T var[5];
  prim_to_cons<T>(T(prim[0]), T(prim[1]), T(prim[2]), T(prim[3]), T(prim[4]), var);
  for (int q = 0; q < 5; ++q) out[q] = (double)var[q];
}

}  // namespace

drnh_grid* drnh_grid_create(int fp64_geometry) {
  drnh_grid* g = new drnh_grid();
  g->fp64 = fp64_geometry != 0;
  if (g->fp64)
    g->gd.reset(new PatchGrid<double>());
  else
    g->gf.reset(new PatchGrid<float>());
  return g;
}

void drnh_grid_destroy(drnh_grid* g) { delete g; }

const char* drnh_last_error(const drnh_grid* g) { return g ? g->err.c_str() : "null grid"; }

int drnh_grid_add_patch(drnh_grid* g, int ni, int nj, int nk, const int32_t* seek6, const double* origin3,
                        const double* length3) {
  if (!g || !seek6 || !origin3 || !length3 || ni < 1 || nj < 1 || nk < 1) return -1;
  int seek[6];
  for (int f = 0; f < 6; ++f) {
    seek[f] = seek6[f];
    if (seek[f] < 0) return -1;
  }
  if (seek[0] + seek[1] >= ni || seek[2] + seek[3] >= nj || seek[4] + seek[5] >= nk) {
    g->err = "seek layers leave no core cells";
    return -1;
  }
  return with(g, [&](auto& G) {
    using T = typename std::remove_reference<decltype(G.patch(0))>::type;
    return G.addPatch(T(ni, nj, nk, seek, origin3, length3));
  });
}

int drnh_grid_read(drnh_grid* g, const char* path, double scale) {
  if (!g || !path) return -1;
  return with(g, [&](auto& G) { return G.readGrid(path, scale, &g->err); });
}

namespace {

int put_string(const std::string& s, char* out, int cap) {
  if (!out || cap < 1) return -1;
  if (s.size() + 1 > (size_t)cap) return -2;
  std::memcpy(out, s.c_str(), s.size() + 1);
  return (int)s.size();
}

}  // namespace

int drnh_codestring(const char* text, char* out, int cap) {
  if (!text) return -1;
  return put_string(CodeString(text), out, cap);
}

int drnh_codestring_compare(const char* a, const char* b, int32_t* flags3, char* out, int cap) {
  if (!a || !b || !flags3) return -1;
  bool same, over, under;
  const CodeString m = CodeString(a).compareCodeString(CodeString(b), same, over, under);
  flags3[0] = same;
  flags3[1] = over;
  flags3[2] = under;
  return put_string(m, out, cap);
}

int drnh_grid_patch_codes(const drnh_grid* g, int p, char* out, int cap) {
  if (bad_patch(g, p)) return -1;
  return with(g, [&](const auto& G) { return put_string(G.patch((size_t)p).solver_codes, out, cap); });
}

int drnh_grid_num_patch_groups(const drnh_grid* g) {
  if (!g) return -1;
  return with(g, [&](const auto& G) { return (int)G.getPatchGroups().accessNumPatchGroups(); });
}

int drnh_grid_patch_group(const drnh_grid* g, int group, int32_t* info3, int32_t* patches, char* out, int cap) {
  if (!g || !info3 || group < 0 || group >= drnh_grid_num_patch_groups(g)) return -1;
  return with(g, [&](const auto& G) {
    const SinglePatchGroup& s = G.getPatchGroups().accessSinglePatchGroup((size_t)group);
    info3[0] = (int32_t)s.patch_type;
    info3[1] = (int32_t)s.patches.size();
    info3[2] = s.solver_exceptions ? 1 : 0;
    if (patches)
      for (size_t i = 0; i < s.patches.size(); ++i) patches[i] = (int32_t)s.patches[i];
    return put_string(s.solver_codes, out, cap);
  });
}

int drnh_grid_num_patches(const drnh_grid* g) {
  if (!g) return -1;
  return with(g, [&](const auto& G) { return (int)G.numPatches(); });
}

int drnh_grid_patch_info(const drnh_grid* g, int p, int32_t* oi, double* od) {
  if (bad_patch(g, p) || !oi || !od) return -1;
  return with(g, [&](const auto& G) {
    const auto& P = G.patch(p);
    oi[0] = P.ni();
    oi[1] = P.nj();
    oi[2] = P.nk();
    for (int f = 0; f < 6; ++f) oi[3 + f] = P.seek(f);
    for (int a = 0; a < 3; ++a) {
      od[a] = (double)P.origin(a);
      od[3 + a] = (double)P.length(a);
      od[6 + a] = (double)P.delta(a);
    }
    return 0;
  });
}

int drnh_grid_set_faces(drnh_grid* g, int p, const int32_t* bc6, const double* par) {
  if (bad_patch(g, p) || !bc6 || !par) return -1;
  for (int f = 0; f < 6; ++f)
    if (bc6[f] < DRN_BC_SLIP || bc6[f] > DRN_BC_NONE) return -1;
  return with(g, [&](auto& G) {
    auto& P = G.patch(p);
    for (int f = 0; f < 6; ++f) {
      P.face[f] = bc6[f];
      for (int q = 0; q < 16; ++q) P.face_par[f][q] = par[16 * f + q];
    }
    return 0;
  });
}

int drnh_grid_compute_dependencies(drnh_grid* g, const double* period3, int threads) {
  if (!g) return -1;
  int r = with(g, [&](auto& G) { return G.computeDependencies(period3, threads); });
  if (r) g->err = "donor search hit an interpolation-range inconsistency";
  return r;
}

int64_t drnh_grid_num_receivers(const drnh_grid* g, int p) {
  if (bad_patch(g, p)) return -1;
  return with(g, [&](const auto& G) { return (int64_t)G.tables(p).receivers.size(); });
}

int64_t drnh_grid_num_no_donor(const drnh_grid* g, int p) {
  if (bad_patch(g, p)) return -1;
  return with(g, [&](const auto& G) { return (int64_t)G.tables(p).no_donor; });
}

int drnh_grid_receivers(const drnh_grid* g, int p, int64_t* out) {
  if (bad_patch(g, p) || !out) return -1;
  return with(g, [&](const auto& G) {
    const auto& v = G.tables(p).receivers;
    for (size_t a = 0; a < v.size(); ++a) out[a] = (int64_t)v[a];
    return 0;
  });
}

int drnh_grid_num_donors(const drnh_grid* g, int p) {
  if (bad_patch(g, p)) return -1;
  return with(g, [&](const auto& G) { return (int)G.tables(p).donors.size(); });
}

int drnh_grid_donor_info(const drnh_grid* g, int p, int d, int64_t* out) {
  if (bad_patch(g, p) || !out) return -1;
  return with(g, [&](const auto& G) {
    const auto& T = G.tables(p);
    if (d < 0 || d >= (int)T.donors.size()) return -1;
    const auto& D = T.donors[d];
    out[0] = D.patch;
    out[1] = (int64_t)D.nrec;
    out[2] = (int64_t)D.stride;
    out[3] = (int64_t)D.rec_start;
    out[4] = (int64_t)D.wi_start;
    return 0;
  });
}

int64_t drnh_grid_num_concat(const drnh_grid* g, int p, int which) {
  if (bad_patch(g, p)) return -1;
  return with(g, [&](const auto& G) {
    const auto& T = G.tables(p);
    return (int64_t)(which == 0 ? T.rec_concat.size() : T.donor_index.size());
  });
}

int drnh_grid_rec_concat(const drnh_grid* g, int p, int64_t* out) {
  if (bad_patch(g, p) || !out) return -1;
  return with(g, [&](const auto& G) {
    const auto& v = G.tables(p).rec_concat;
    for (size_t a = 0; a < v.size(); ++a) out[a] = (int64_t)v[a];
    return 0;
  });
}

int drnh_grid_donor_index(const drnh_grid* g, int p, int64_t* out) {
  if (bad_patch(g, p) || !out) return -1;
  return with(g, [&](const auto& G) {
    const auto& v = G.tables(p).donor_index;
    for (size_t a = 0; a < v.size(); ++a) out[a] = (int64_t)v[a];
    return 0;
  });
}

int drnh_grid_donor_weight(const drnh_grid* g, int p, double* out) {
  if (bad_patch(g, p) || !out) return -1;
  return with(g, [&](const auto& G) {
    const auto& v = G.tables(p).donor_weight;
    if (!v.empty()) memcpy(out, v.data(), v.size() * sizeof(double));
    return 0;
  });
}

int drnh_grid_donor_turn(const drnh_grid* g, int p, int d, double* out) {
  if (bad_patch(g, p) || !out) return -1;
  return with(g, [&](const auto& G) {
    const auto& D = G.tables(p).donors;
    if (d < 0 || d >= (int)D.size()) return -1;
    memcpy(out, D[(size_t)d].turn, 9 * sizeof(double));
    return 0;
  });
}

int drnh_grid_set_base(drnh_grid* g, int p, const double* bi, const double* bj) {
  if (bad_patch(g, p) || !bi || !bj) return -1;
  const double cx = bi[1] * bj[2] - bi[2] * bj[1], cy = bi[2] * bj[0] - bi[0] * bj[2], cz = bi[0] * bj[1] - bi[1] * bj[0];
  const double ni2 = bi[0] * bi[0] + bi[1] * bi[1] + bi[2] * bi[2], nj2 = bj[0] * bj[0] + bj[1] * bj[1] + bj[2] * bj[2];
  if (!(ni2 > 0) || !(nj2 > 0) || !(cx * cx + cy * cy + cz * cz > 1e-12 * ni2 * nj2)) {
    g->err = "patch base vectors are zero or parallel";
    return -1;
  }
  return with(g, [&](auto& G) {
    G.patch(p).setBase(bi, bj);
    return 0;
  });
}

int drnh_grid_patch_base(const drnh_grid* g, int p, double* out) {
  if (bad_patch(g, p) || !out) return -1;
  return with(g, [&](const auto& G) {
    const auto& P = G.patch(p);
    for (int a = 0; a < 3; ++a) {
      out[a] = (double)P.baseI(a);
      out[3 + a] = (double)P.baseJ(a);
    }
    return 0;
  });
}

int drnh_grid_patch_transform(const drnh_grid* g, int p, double* out) {
  if (bad_patch(g, p) || !out) return -1;
  return with(g, [&](const auto& G) {
    const auto& t = G.patch(p).this2inertial();
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) out[3 * r + c] = (double)t.a[r][c];
      out[9 + r] = (double)t.b[r];
    }
    return 0;
  });
}

int drnh_grid_levelset_spheres(drnh_grid* g, int n_spheres, const double* spheres, int n_inner, int n_outer,
                               double min_inner_rel) {
  if (!g || n_spheres < 1 || !spheres || n_inner < 0 || n_outer < 0) return -1;
  if (min_inner_rel > 0 && n_inner == 0) return -1;
  std::vector<double> s(spheres, spheres + 4 * (size_t)n_spheres);
  int r = with(g, [&](auto& G) { return G.levelSetSpheres(s, n_inner, n_outer, min_inner_rel); });
  if (r) g->err = "mirror-point interpolation hit an interpolation-range inconsistency";
  return r;
}

int drnh_grid_levelset_values(const drnh_grid* g, int p, double* out) {
  if (bad_patch(g, p) || !out) return -1;
  return with(g, [&](const auto& G) {
    const auto& v = G.levelSetValues(p);
    for (size_t a = 0; a < v.size(); ++a) out[a] = (double)v[a];
    return 0;
  });
}

int64_t drnh_grid_ls_num_cells(const drnh_grid* g, int p, int layer, int outer) {
  if (bad_patch(g, p)) return -1;
  return with(g, [&](const auto& G) -> int64_t {
    if (outer) return layer >= 0 && layer < G.numOuterLayers() ? (int64_t)G.outerLayer(p, layer).size() : 0;
    return layer >= 0 && layer < G.numInnerLayers() ? (int64_t)G.innerLayer(p, layer).size() : 0;
  });
}

int drnh_grid_ls_cells(const drnh_grid* g, int p, int layer, int outer, int64_t* cells, double* gv, int64_t* donors,
                       double* weights) {
  if (bad_patch(g, p)) return -1;
  return with(g, [&](const auto& G) {
    const int nl = outer ? G.numOuterLayers() : G.numInnerLayers();
    if (layer < 0 || layer >= nl) return -1;
    const auto& L = outer ? G.outerLayer(p, layer) : G.innerLayer(p, layer);
    for (size_t a = 0; a < L.size(); ++a) {
      if (cells) cells[a] = (int64_t)L[a].cell;
      if (gv) {
        gv[4 * a] = L[a].g;
        gv[4 * a + 1] = L[a].gx;
        gv[4 * a + 2] = L[a].gy;
        gv[4 * a + 3] = L[a].gz;
      }
      for (int s = 0; s < 8; ++s) {
        if (donors) donors[8 * a + s] = (int64_t)L[a].donor[s];
        if (weights) weights[8 * a + s] = L[a].weight[s];
      }
    }
    return 0;
  });
}

double drnh_grid_min_ch_length(const drnh_grid* g) {
  if (!g) return -1;
  return with(g, [&](const auto& G) { return G.minChLength(); });
}

int drnh_grid_to_device(drnh_grid* g, drn_ctx* ctx) {
  if (!g || !ctx) return -1;
  const int prec = drn_precision(ctx);
  return with(g, [&](auto& G) {
    for (size_t p = 0; p < G.numPatches(); ++p) {
      const auto& P = G.patch(p);
      drn_patch_desc d;
      memset(&d, 0, sizeof(d));
      d.ni = P.ni();
      d.nj = P.nj();
      d.nk = P.nk();
      for (int f = 0; f < 6; ++f) d.seek[f] = P.seek(f);
      // CartesianPatch::computeDeltas in the state precision (m_Dx = m_Lx/m_NumI)
      double dd[3];
      const int n[3] = {P.ni(), P.nj(), P.nk()};
      for (int a = 0; a < 3; ++a)
        dd[a] = prec == DRN_F32 ? (double)((float)P.length(a) / (float)(size_t)n[a])
                                : (double)P.length(a) / (double)(size_t)n[a];
      d.dx = dd[0];
      d.dy = dd[1];
      d.dz = dd[2];
      for (int f = 0; f < 6; ++f) {
        d.bc[f] = P.face[f];
        for (int q = 0; q < 16; ++q) d.bc_par[f][q] = P.face_par[f][q];
        if (P.face[f] == FACE_FARFIELD) {
          if (prec == DRN_F32)
            farfield_to_conserved<float>(P.face_par[f], d.bc_par[f]);
          else
            farfield_to_conserved<double>(P.face_par[f], d.bc_par[f]);
        }
      }
      int id = -1;
      int rc = drn_patch_add(ctx, &d, &id);
      if (rc) {
        g->err = std::string("drn_patch_add: ") + drn_last_error(ctx);
        return rc;
      }
    }
    int rc = drn_layout(ctx);
    if (rc) g->err = std::string("drn_layout: ") + drn_last_error(ctx);
    return rc;
  });
}

int drnh_grid_upload_exchange(drnh_grid* g, drn_ctx* ctx) {
  if (!g || !ctx) return -1;
  return with(g, [&](auto& G) {
    for (size_t p = 0; p < G.numPatches(); ++p) {
      const DonorTables& T = G.tables(p);
      if (T.donors.empty()) continue;
      std::vector<int32_t> dp;
      std::vector<int64_t> info;
      for (const auto& D : T.donors) {
        dp.push_back(D.patch);
        info.push_back((int64_t)D.nrec);
        info.push_back((int64_t)D.stride);
        info.push_back((int64_t)D.rec_start);
        info.push_back((int64_t)D.wi_start);
      }
      std::vector<int64_t> ru(T.receivers.begin(), T.receivers.end());
      std::vector<int64_t> rc(T.rec_concat.begin(), T.rec_concat.end());
      std::vector<int64_t> di(T.donor_index.begin(), T.donor_index.end());
      int r = drn_exchange_set(ctx, (int)p, (int)dp.size(), dp.data(), info.data(), (int64_t)ru.size(), ru.data(),
                               (int64_t)rc.size(), rc.data(), (int64_t)di.size(), di.data(), T.donor_weight.data());
      if (r) {
        g->err = std::string("drn_exchange_set: ") + drn_last_error(ctx);
        return r;
      }
      // the momentum's turn, where a donor is not aligned with this patch
      std::vector<double> turn;
      bool aligned = true;
      for (const auto& D : T.donors) {
        turn.insert(turn.end(), D.turn, D.turn + 9);
        for (int e = 0; e < 9; ++e) aligned = aligned && D.turn[e] == (e % 4 == 0 ? 1.0 : 0.0);
      }
      if (!aligned) {
        r = drn_exchange_set_turn(ctx, (int)p, (int)T.donors.size(), turn.data());
        if (r) {
          g->err = std::string("drn_exchange_set_turn: ") + drn_last_error(ctx);
          return r;
        }
      }
    }
    int r = drn_exchange_commit(ctx);
    if (r) g->err = std::string("drn_exchange_commit: ") + drn_last_error(ctx);
    return r;
  });
}

int drnh_grid_upload_levelset(drnh_grid* g, drn_ctx* ctx) {
  if (!g || !ctx) return -1;
  return with(g, [&](auto& G) {
    for (size_t p = 0; p < G.numPatches(); ++p) {
      std::vector<int64_t> cells, donors;
      std::vector<double> nrm, w;
      for (int layer = 0; layer < G.numInnerLayers(); ++layer)
        for (const LayerCell& c : G.innerLayer(p, layer)) {
          cells.push_back((int64_t)c.cell);
          nrm.push_back(c.gx);
          nrm.push_back(c.gy);
          nrm.push_back(c.gz);
          for (int s = 0; s < 8; ++s) {
            donors.push_back((int64_t)c.donor[s]);
            w.push_back(c.weight[s]);
          }
        }
      if (cells.empty()) continue;
      int r = drn_levelset_set(ctx, (int)p, (int64_t)cells.size(), cells.data(), nrm.data(), donors.data(), w.data());
      if (r) {
        g->err = std::string("drn_levelset_set: ") + drn_last_error(ctx);
        return r;
      }
    }
    int r = drn_levelset_commit(ctx);
    if (r) g->err = std::string("drn_levelset_commit: ") + drn_last_error(ctx);
    return r;
  });
}
