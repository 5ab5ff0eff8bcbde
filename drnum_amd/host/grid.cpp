// grid.cpp -- CartesianPatch geometry, donor tables and level-set layers.
//
// The expressions keep the literal types of the reference (a double literal in
// a `real` expression promotes it to double), so that G = float reproduces the
// reference's float build and G = double its FP64 branch bit for bit.
#include "grid.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iterator>
#include <sstream>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace drnum {

// ---------------------------------------------------------------------------
// WeightSet
// ---------------------------------------------------------------------------

template <typename G>
void WeightSet<G>::unify() {
  // sort by (index, weight) as std::pair ordering does, then add the weights of equal indices
  for (int a = 1; a < n; ++a) {
    size_t ti = idx[a];
    G tw = w[a];
    int b = a;
    while (b > 0 && (idx[b - 1] > ti || (idx[b - 1] == ti && w[b - 1] > tw))) {
      idx[b] = idx[b - 1];
      w[b] = w[b - 1];
      --b;
    }
    idx[b] = ti;
    w[b] = tw;
  }
  if (n == 0) return;
  int hold = 0;
  for (int it = 1; it < n; ++it) {
    if (idx[it] == idx[hold]) {
      w[hold] += w[it];
    } else {
      ++hold;
      idx[hold] = idx[it];
      w[hold] = w[it];
    }
  }
  n = hold + 1;
}

template <typename G>
void WeightSet<G>::eliminateBelow(G eps) {
  unify();
  int m = 0;
  for (int a = 0; a < n; ++a) {
    G t = w[a];
    if (t < G(0)) t = -t;
    if (!(t < eps)) {
      idx[m] = idx[a];
      w[m] = w[a];
      ++m;
    }
  }
  n = m;
}

template <typename G>
void WeightSet<G>::shiftToSum(G sum) {
  if (n == 0) return;
  G old = 0;
  for (int a = 0; a < n; ++a) old += w[a];
  G corr = (sum - old) / G(size_t(n));
  for (int a = 0; a < n; ++a) w[a] += corr;
}

// ---------------------------------------------------------------------------
// Transform
// ---------------------------------------------------------------------------

template <typename G>
Transform<G> Transform<G>::identity() {
  Transform t;
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) t.a[r][c] = r == c ? G(1) : G(0);
    t.b[r] = 0;
  }
  return t;
}

namespace {

// MathVector::abs and normalise (mathvector_methods.h:122-127, 178-183)
template <typename G>
G abs3(const G* v) {
  G l = 0;
  for (int i = 0; i < 3; ++i) l += v[i] * v[i];
  return std::sqrt(l);
}
template <typename G>
void normalise3(G* v) {
  const G l = abs3(v);
  for (int i = 0; i < 3; ++i) v[i] /= l;
}

}  // namespace

template <typename G>
void Transform<G>::setBaseIJ(const G* base_i, const G* base_j) {
  G bi[3] = {base_i[0], base_i[1], base_i[2]}, bj[3] = {base_j[0], base_j[1], base_j[2]}, bk[3];
  normalise3(bi);
  normalise3(bj);
  // cross product (mathvector_methods.h:86-93)
  bk[0] = bi[1] * bj[2] - bi[2] * bj[1];
  bk[1] = bi[2] * bj[0] - bi[0] * bj[2];
  bk[2] = bi[0] * bj[1] - bi[1] * bj[0];
  const G bk_len = abs3(bk);
  G scal_ij = 0;  // scalar product (mathvector_operators.h:208-216)
  for (int i = 0; i < 3; ++i) scal_ij += bi[i] * bj[i];
  for (int i = 0; i < 3; ++i) bj[i] = bj[i] - scal_ij * bi[i];
  normalise3(bj);
  // (1./len) * base_k: the double quotient becomes the `real` factor of operator*
  const G inv_len = G(1. / bk_len);
  for (int i = 0; i < 3; ++i) bk[i] = inv_len * bk[i];
  for (int i = 0; i < 3; ++i) {
    a[i][0] = bi[i];
    a[i][1] = bj[i];
    a[i][2] = bk[i];
  }
}

template <typename G>
Transform<G> Transform<G>::inverse() const {
  // InvSmallSquareMatrix<T,3> (smallsquarematrix.h:470-489), the branch of a
  // matrix without precision checks, then the shift -A^-1 b as M*v
  // (smallsquarematrix.h:209-218)
  const G(*M)[3] = a;
  const G t4 = M[0][0] * M[1][1];
  const G t6 = M[0][0] * M[1][2];
  const G t8 = M[0][1] * M[1][0];
  const G t10 = M[0][2] * M[1][0];
  const G t12 = M[0][1] * M[2][0];
  const G t14 = M[0][2] * M[2][0];
  const G t17 = 1 / (t4 * M[2][2] - t6 * M[2][1] - t8 * M[2][2] + t10 * M[2][1] + t12 * M[1][2] - t14 * M[1][1]);
  Transform t;
  t.a[0][0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) * t17;
  t.a[0][1] = -(M[0][1] * M[2][2] - M[0][2] * M[2][1]) * t17;
  t.a[0][2] = -(-M[0][1] * M[1][2] + M[0][2] * M[1][1]) * t17;
  t.a[1][0] = -(M[1][0] * M[2][2] - M[1][2] * M[2][0]) * t17;
  t.a[1][1] = (M[0][0] * M[2][2] - t14) * t17;
  t.a[1][2] = -(t6 - t10) * t17;
  t.a[2][0] = -(-M[1][0] * M[2][1] + M[1][1] * M[2][0]) * t17;
  t.a[2][1] = -(M[0][0] * M[2][1] - t12) * t17;
  t.a[2][2] = (t4 - t8) * t17;
  const G v[3] = {-b[0], -b[1], -b[2]};
  for (int i = 0; i < 3; ++i) {
    t.b[i] = 0;
    for (int j = 0; j < 3; ++j) t.b[i] += t.a[i][j] * v[j];
  }
  return t;
}

template <typename G>
Transform<G> Transform<G>::concatenate(const Transform& c) const {
  Transform t;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) t.a[i][j] = a[i][0] * c.a[0][j] + a[i][1] * c.a[1][j] + a[i][2] * c.a[2][j];
    t.b[i] = a[i][0] * c.b[0] + a[i][1] * c.b[1] + a[i][2] * c.b[2] + b[i];
  }
  return t;
}

template <typename G>
void Transform<G>::transform(G x, G y, G z, G* out) const {
  for (int i = 0; i < 3; ++i) out[i] = a[i][0] * x + a[i][1] * y + a[i][2] * z + b[i];
}

template <typename G>
void setupTransformation(const G* origin, const G* base_i, const G* base_j, Transform<G>& i2t, Transform<G>& t2i) {
  Transform<G> this2inertial = Transform<G>::identity();
  for (int a = 0; a < 3; ++a) this2inertial.b[a] = origin[a];
  this2inertial.setBaseIJ(base_i, base_j);
  i2t = this2inertial.inverse();
  t2i = i2t.inverse();
}

// ---------------------------------------------------------------------------
// CartesianPatch
// ---------------------------------------------------------------------------

template <typename G>
CartesianPatch<G>::CartesianPatch(int ni, int nj, int nk, const int* seek, const double* origin,
                                  const double* length, const double* base_i, const double* base_j) {
  n_[0] = ni;
  n_[1] = nj;
  n_[2] = nk;
  njk_ = (size_t)nj * nk;
  for (int f = 0; f < 6; ++f) seek_[f] = seek[f];
  for (int a = 0; a < 3; ++a) {
    o_[a] = G(origin[a]);
    L_[a] = G(length[a]);
    // computeDeltas (cartesianpatch.cpp:164-212): m_Dx = m_Lx/m_NumI, m_xCCMin = 0.5*m_Dx
    d_[a] = L_[a] / G(size_t(n_[a]));
    ccmin_[a] = 0.5 * d_[a];
    // buildRegions (cartesianpatch.cpp:314-390) without additional protection layers
    dz_first_[a] = (size_t)seek_[2 * a];
    dz_after_[a] = (size_t)n_[a] - (size_t)seek_[2 * a + 1];
    bi_[a] = base_i ? G(base_i[a]) : G(a == 0 ? 1 : 0);
    bj_[a] = base_j ? G(base_j[a]) : G(a == 1 ? 1 : 0);
  }
  setupTransformation(o_, bi_, bj_, i2t_, t2i_);
}

template <typename G>
void CartesianPatch<G>::setBase(const double* base_i, const double* base_j) {
  for (int a = 0; a < 3; ++a) {
    bi_[a] = G(base_i[a]);
    bj_[a] = G(base_j[a]);
  }
  setupTransformation(o_, bi_, bj_, i2t_, t2i_);
}

template <typename G>
void CartesianPatch<G>::xyzCell(size_t l, G& x, G& y, G& z) const {
  size_t i, j, k;
  ijk(l, i, j, k);
  x = ccmin_[0] + G(i) * d_[0];
  y = ccmin_[1] + G(j) * d_[1];
  z = ccmin_[2] + G(k) * d_[2];
}

template <typename G>
bool CartesianPatch<G>::isCore(size_t i, size_t j, size_t k) const {
  return i >= (size_t)seek_[0] && i < (size_t)(n_[0] - seek_[1]) && j >= (size_t)seek_[2] &&
         j < (size_t)(n_[1] - seek_[3]) && k >= (size_t)seek_[4] && k < (size_t)(n_[2] - seek_[5]);
}

template <typename G>
std::vector<size_t> CartesianPatch<G>::seekCells() const {
  std::vector<size_t> r;
  for (size_t i = 0; i < (size_t)n_[0]; ++i)
    for (size_t j = 0; j < (size_t)n_[1]; ++j)
      for (size_t k = 0; k < (size_t)n_[2]; ++k)
        if (!isCore(i, j, k)) r.push_back(index(i, j, k));
  return r;
}

// linearCCInterpolCoeffs (cartesianpatch.cpp:1053-1162)
template <typename G>
bool CartesianPatch<G>::donorZoneMayContain(const double* lo, const double* hi) const {
  for (int a = 0; a < 3; ++a) {
    const double d = (double)d_[a];
    const double z0 = (double)dz_first_[a] * d, z1 = (double)dz_after_[a] * d;
    const double tol = 1e-6 * (std::fabs(z0) + std::fabs(z1) + std::fabs(lo[a]) + std::fabs(hi[a]) + d);
    if (hi[a] < z0 - tol || lo[a] > z1 + tol) return false;
  }
  return true;
}

template <typename G>
void CartesianPatch<G>::cellCentreBox(double* lo, double* hi) const {
  for (int a = 0; a < 3; ++a) {
    lo[a] = (double)ccmin_[a];
    hi[a] = (double)ccmin_[a] + (double)(n_[a] - 1) * (double)d_[a];
  }
}

template <typename G>
int CartesianPatch<G>::lin(G s_test, int axis, size_t& ref, G& coeff) const {
  const G m_eps = 1.e-5;  // cartesianpatch.cpp:40
  const size_t first = dz_first_[axis], afterlast = dz_after_[axis], num = (size_t)n_[axis];
  const G ds = d_[axis];
  G rs_test = s_test / ds;
  G rs0 = float(first);
  G rs1 = rs0 + 0.5;
  G rs2 = float(afterlast) - 0.5;
  G rs3 = rs2 + 0.5;
  int sector;
  if (rs_test < rs0) return 0;
  if (rs_test < rs1)
    sector = 1;
  else if (rs_test < rs2)
    sector = 2;
  else if (rs_test < rs3)
    sector = 3;
  else
    return 0;
  if (sector == 1) {
    ref = first;
    coeff = 0.;
  } else if (sector == 2) {
    long l = long(rs_test - 0.5 + float(num)) - (long)num;
    if ((size_t)l == first - 1) {
      l = (long)first;
    } else if ((size_t)l == afterlast - 1) {
      l = (long)afterlast - 2;
    } else if (l < -1 || (size_t)l > afterlast - 1) {
      return -1;
    }
    ref = size_t(l);
    G rs_cell_ref = 0.5 + float(ref);
    coeff = rs_test - rs_cell_ref;
  } else {
    ref = afterlast - 2;
    coeff = 1.;
  }
  if (ref < first || ref > afterlast - 2) return -1;
  if (coeff < (-m_eps)) return -1;
  if (coeff > (1. + m_eps)) return -1;
  return 1;
}

template <typename G>
int CartesianPatch<G>::interpolCoeffs(G x, G y, G z, WeightSet<G>& ws) const {
  size_t ic, jc, kc;
  G xc, yc, zc;
  int r;
  if ((r = lin(x, 0, ic, xc)) != 1) return r;
  if ((r = lin(y, 1, jc, yc)) != 1) return r;
  if ((r = lin(z, 2, kc, zc)) != 1) return r;
  G w[8];
  w[0] = (1. - xc) * (1. - yc) * (1. - zc);
  w[1] = (1. - xc) * (1. - yc) * zc;
  w[2] = (1. - xc) * yc * (1. - zc);
  w[3] = (1. - xc) * yc * zc;
  w[4] = xc * (1. - yc) * (1. - zc);
  w[5] = xc * (1. - yc) * zc;
  w[6] = xc * yc * (1. - zc);
  w[7] = xc * yc * zc;
  size_t ic1 = ic + 1, jc1 = jc + 1, kc1 = kc + 1;
  if (ic1 + 1 > (size_t)n_[0]) ic1 = ic;
  if (jc1 + 1 > (size_t)n_[1]) jc1 = jc;
  if (kc1 + 1 > (size_t)n_[2]) kc1 = kc;
  const size_t cid[8] = {index(ic, jc, kc),  index(ic, jc, kc1),  index(ic, jc1, kc),  index(ic, jc1, kc1),
                         index(ic1, jc, kc), index(ic1, jc, kc1), index(ic1, jc1, kc), index(ic1, jc1, kc1)};
  ws.n = 8;
  for (int a = 0; a < 8; ++a) {
    ws.idx[a] = cid[a];
    ws.w[a] = w[a];
  }
  const G m_eps = 1.e-5;
  ws.eliminateBelow(10 * m_eps);
  ws.shiftToSum(1.);
  return 1;
}

// ---------------------------------------------------------------------------
// PatchGrid
// ---------------------------------------------------------------------------

template <typename G>
int PatchGrid<G>::addPatch(const CartesianPatch<G>& p) {
  patches_.push_back(p);
  return (int)patches_.size() - 1;
}

template <typename G>
int PatchGrid<G>::readGrid(const std::string& path, double scale, std::string* err) {
  std::ifstream f(path);
  if (!f) {
    if (err) *err = "cannot open grid file " + path;
    return -1;
  }
  while (f.good()) {
    size_t type = 0;
    f >> type;
    if (!f || type == 0) break;
    std::string comment;
    char c;
    while (f.get(c) && c != '{') comment.push_back(c);
    std::string body;
    std::getline(f, body, '}');
    for (auto& ch : body)
      if (isspace((unsigned char)ch)) ch = ' ';
    if (type != 1001) {
      if (err) *err = "unsupported patch type " + std::to_string(type);
      return -1;
    }
    // Patch::readFromFile (patch.cpp:63-85) + CartesianPatch::readFromFile (cartesianpatch.cpp:98-133)
    // values are read into `real` (G) directly, as the reference's stream extraction does
    std::istringstream is(body);
    G o[3], bi[3], bj[3], ioscale, L[3];
    int n[3], s[6];
    is >> o[0] >> o[1] >> o[2] >> bi[0] >> bi[1] >> bi[2] >> bj[0] >> bj[1] >> bj[2] >> ioscale;
    is >> n[0] >> n[1] >> n[2];
    for (int a = 0; a < 6; ++a) is >> s[a];
    is >> L[0] >> L[1] >> L[2];
    if (!is) {
      if (err) *err = "malformed patch entry in " + path;
      return -1;
    }
    {
      // setMatrixFromBaseIJ divides by |base_i|, |base_j| and |base_i x base_j|
      const double bid[3] = {(double)bi[0], (double)bi[1], (double)bi[2]};
      const double bjd[3] = {(double)bj[0], (double)bj[1], (double)bj[2]};
      const double cx = bid[1] * bjd[2] - bid[2] * bjd[1], cy = bid[2] * bjd[0] - bid[0] * bjd[2],
                   cz = bid[0] * bjd[1] - bid[1] * bjd[0];
      const double ni2 = bid[0] * bid[0] + bid[1] * bid[1] + bid[2] * bid[2];
      const double nj2 = bjd[0] * bjd[0] + bjd[1] * bjd[1] + bjd[2] * bjd[2];
      if (!(ni2 > 0) || !(nj2 > 0) || !(cx * cx + cy * cy + cz * cz > 1e-12 * ni2 * nj2)) {
        if (err) *err = "patch base vectors are zero or parallel in " + path;
        return -1;
      }
    }
    ioscale *= G(scale);
    double od[3], Ld[3], bid[3], bjd[3];
    for (int a = 0; a < 3; ++a) {
      o[a] *= ioscale;
      L[a] *= ioscale;
      od[a] = (double)o[a];
      Ld[a] = (double)L[a];
      bid[a] = (double)bi[a];
      bjd[a] = (double)bj[a];
    }
    CartesianPatch<G> P(n[0], n[1], n[2], s, od, Ld, bid, bjd);
    P.comment = comment;
    // the rest of the entry is the patch's code string (CodeString::buildFrom)
    P.solver_codes = CodeString(std::string(std::istreambuf_iterator<char>(is), std::istreambuf_iterator<char>()));
    const int id = addPatch(P);
    groups_.insertPatch((size_t)id, 1001, patches_[(size_t)id].solver_codes);
  }
  return 0;
}

template <typename G>
double PatchGrid<G>::minChLength() const {
  double m = 0;
  for (size_t p = 0; p < patches_.size(); ++p) {
    double v = std::min(std::min((double)patches_[p].delta(0), (double)patches_[p].delta(1)),
                        (double)patches_[p].delta(2));
    if (p == 0 || v < m) m = v;
  }
  return m;
}

namespace {

template <typename G>
struct Contribution {
  size_t ll, l;
  WeightSet<G> ws;
};

struct Turn {
  double a[9];
};

}  // namespace

template <typename G>
int PatchGrid<G>::computeDependencies(const double* period, int threads) {
  const int np = (int)patches_.size();
  tables_.assign(np, DonorTables());
  // periodic image shifts, in the order (-1,0,1)^3 without the zero shift
  std::vector<std::array<int, 3>> shifts;
  for (int a = -1; a <= 1; ++a)
    for (int b = -1; b <= 1; ++b)
      for (int c = -1; c <= 1; ++c) {
        const int s[3] = {a, b, c};
        bool ok = !(a == 0 && b == 0 && c == 0);
        for (int x = 0; x < 3; ++x)
          if (s[x] != 0 && !(period && period[x] > 0)) ok = false;
        if (ok) shifts.push_back({a, b, c});
      }
  int err = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1) reduction(| : err)
#endif
  for (int p = 0; p < np; ++p) {
    const CartesianPatch<G>& P = patches_[p];
    DonorTables& T = tables_[p];
    std::vector<size_t> recv = P.seekCells();
    if (recv.empty()) continue;
    std::vector<size_t> hits(recv.size(), 0);
    std::vector<std::pair<int, std::vector<Contribution<G>>>> neigh;
    std::vector<Turn> turns;  // per entry of neigh
    // Patch::insertNeighbour (patch.cpp:264-317) for every candidate donor in ascending order
    for (int q = 0; q < np && !err; ++q) {
      for (int s = -1; s < (int)shifts.size(); ++s) {
        if (s < 0 && q == p) continue;
        const CartesianPatch<G>& Q = patches_[q];
        G oq[3];
        for (int a = 0; a < 3; ++a)
          oq[a] = (s < 0 || shifts[s][a] == 0) ? Q.origin(a) : Q.origin(a) + G(shifts[s][a] * period[a]);
        // CoordTransformVV::setTransFromTo(receiver, donor) (coordtransformvv.cpp:65-71,
        // patch.cpp:285-294): `rel` takes receiver cell centres into the donor's
        // frame; `turn` (its stored inverse) turns the donor's momentum
        Transform<G> q_i2t, q_t2i;
        setupTransformation(oq, Q.baseI(), Q.baseJ(), q_i2t, q_t2i);
        const Transform<G> rel = P.this2inertial().concatenate(q_i2t);
        const Transform<G> turn = q_t2i.concatenate(P.inertial2this());
        {
          // the receiver's cell-centre box in the donor frame, widened to the
          // box of its eight transformed corners
          double lo[3], hi[3], blo[3] = {1e300, 1e300, 1e300}, bhi[3] = {-1e300, -1e300, -1e300};
          P.cellCentreBox(lo, hi);
          for (int c = 0; c < 8; ++c) {
            const double q[3] = {(c & 1) ? hi[0] : lo[0], (c & 2) ? hi[1] : lo[1], (c & 4) ? hi[2] : lo[2]};
            for (int r = 0; r < 3; ++r) {
              const double v = (double)rel.a[r][0] * q[0] + (double)rel.a[r][1] * q[1] + (double)rel.a[r][2] * q[2] +
                               (double)rel.b[r];
              blo[r] = std::min(blo[r], v);
              bhi[r] = std::max(bhi[r], v);
            }
          }
          if (!Q.donorZoneMayContain(blo, bhi)) continue;
        }
        std::vector<Contribution<G>> list;
        for (size_t ll = 0; ll < recv.size(); ++ll) {
          G x, y, z;
          P.xyzCell(recv[ll], x, y, z);
          G xt[3];
          rel.transform(x, y, z, xt);
          WeightSet<G> ws;
          int r = Q.interpolCoeffs(xt[0], xt[1], xt[2], ws);
          if (r < 0) {
            err = 1;
            break;
          }
          if (r == 1) {
            list.push_back({ll, recv[ll], ws});
            hits[ll]++;
          }
        }
        if (!list.empty()) {
          Turn tn;
          for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) tn.a[3 * r + c] = (double)turn.a[r][c];
          neigh.push_back({q, std::move(list)});
          turns.push_back(tn);
        }
      }
    }
    if (err) continue;
    // Patch::finalizeDependencies (patch.cpp:319-465)
    std::vector<size_t> new_from_old(recv.size());
    size_t shift = 0;
    for (size_t ll = 0; ll < recv.size(); ++ll) {
      new_from_old[ll] = ll - shift;
      if (hits[ll] == 0) ++shift;
    }
    T.no_donor = shift;
    size_t m = 0;
    for (size_t ll = 0; ll < recv.size(); ++ll)
      if (hits[ll] > 0) {
        recv[m] = recv[ll];
        hits[m] = hits[ll];
        ++m;
      }
    recv.resize(m);
    hits.resize(m);
    T.receivers = recv;
    size_t nrc = 0, nwi = 0;
    for (auto& nb : neigh) {
      size_t stride = 0;
      for (auto& e : nb.second) {
        e.ll = new_from_old[e.ll];
        // InterCoeffWS::adjust2Average (intercoeffws.h:233-242)
        if (hits[e.ll] > 1) {
          const G sc = (1. / hits[e.ll]);
          for (int a = 0; a < e.ws.n; ++a) e.ws.w[a] *= sc;
        }
        stride = std::max(stride, (size_t)e.ws.n);
      }
      DonorTables::Donor D{nb.first, nb.second.size(), stride, nrc, nwi, {}};
      std::memcpy(D.turn, turns[T.donors.size()].a, sizeof(D.turn));
      T.donors.push_back(D);
      nrc += nb.second.size();
      nwi += nb.second.size() * stride;
    }
    // InterCoeffPad::build (intercoeffpad.h:139-170) + buildDonorTransferData (patch.cpp:128-218)
    T.rec_concat.resize(nrc);
    T.donor_index.resize(nwi);
    T.donor_weight.resize(nwi);
    for (size_t d = 0; d < neigh.size(); ++d) {
      const auto& D = T.donors[d];
      for (size_t e = 0; e < neigh[d].second.size(); ++e) {
        const Contribution<G>& c = neigh[d].second[e];
        T.rec_concat[D.rec_start + e] = c.l;
        const size_t st = D.wi_start + e * D.stride;
        for (size_t a = 0; a < D.stride; ++a) {
          if ((int)a < c.ws.n) {
            T.donor_index[st + a] = c.ws.idx[a];
            T.donor_weight[st + a] = (double)c.ws.w[a];
          } else {
            T.donor_index[st + a] = T.donor_index[st];
            T.donor_weight[st + a] = 0.;
          }
        }
      }
    }
  }
  return err ? -1 : 0;
}

// cartesianpatch.cpp:1512-1549, order k-, k+, j-, j+, i-, i+
template <typename G>
static int faceNeighbours(const CartesianPatch<G>& P, size_t l, size_t* nb) {
  size_t i, j, k;
  P.ijk(l, i, j, k);
  const size_t c[6][3] = {{i, j, k - 1}, {i, j, k + 1}, {i, j - 1, k}, {i, j + 1, k}, {i - 1, j, k}, {i + 1, j, k}};
  int n = 0;
  for (int a = 0; a < 6; ++a)
    if (P.inRange(c[a][0], c[a][1], c[a][2])) nb[n++] = P.index(c[a][0], c[a][1], c[a][2]);
  return n;
}

template <typename G>
int PatchGrid<G>::levelSetSpheres(const std::vector<double>& spheres, int n_inner, int n_outer,
                                  double min_inner_rel) {
  const int np = (int)patches_.size();
  n_inner_ = n_inner;
  n_outer_ = n_outer;
  lsg_.assign(np, {});
  inner_.assign(np, std::vector<std::vector<LayerCell>>(std::max(n_inner, 0)));
  outer_.assign(np, std::vector<std::vector<LayerCell>>(std::max(n_outer, 0)));
  const size_t ns = spheres.size() / 4;
  for (int p = 0; p < np; ++p) {
    const CartesianPatch<G>& P = patches_[p];
    const size_t N = P.cells();
    std::vector<G>& var = lsg_[p];
    var.resize(N);
    // LevelSetObject::update: G at the cell centres in the inertial frame
    for (size_t l = 0; l < N; ++l) {
      G x, y, z;
      P.xyzCell(l, x, y, z);
      // Patch::xyzoCell (patch.cpp:91-98): m_TransformInertial2This.transformReverse
      G xyzo[3];
      P.this2inertial().transform(x, y, z, xyzo);
      const G xo = xyzo[0], yo = xyzo[1], zo = xyzo[2];
      G dist = 0;
      for (size_t s = 0; s < ns; ++s) {
        const G dx = xo - G(spheres[4 * s]), dy = yo - G(spheres[4 * s + 1]), dz = zo - G(spheres[4 * s + 2]);
        const G q = dx * dx + dy * dy + dz * dz;
        const G dh = std::sqrt(q) - G(spheres[4 * s + 3]);
        if (s == 0 || dist > dh) dist = dh;
      }
      var[l] = dist;
    }
    std::vector<char> marker(N, 0);
    G gmax = var[0], gmin = var[0];
    for (size_t l = 0; l < N; ++l) {
      if (gmin > var[l]) gmin = var[l];
      if (gmax < var[l]) gmax = var[l];
    }
    const G g_eps = (gmax - gmin) * 1.e-7;
    auto mk = [&](size_t l) {
      LayerCell c{};
      c.cell = l;
      c.g = (double)var[l];
      return c;
    };
    size_t nb[6];
    for (size_t l = 0; l < N; ++l) {
      const G g = var[l];
      const int n = faceNeighbours(P, l, nb);
      bool other = false;
      for (int a = 0; a < n; ++a) {
        const G t1 = (var[nb[a]] + g_eps) * (g - g_eps);
        const G t2 = (var[nb[a]] - g_eps) * (g + g_eps);
        if (t1 < 0. || t2 < 0.) {
          other = true;
          marker[l] = 1;
          break;
        }
      }
      if (other) {
        if (g < 0.) {
          if (n_inner > 0) inner_[p][0].push_back(mk(l));
        } else {
          if (n_outer > 0) outer_[p][0].push_back(mk(l));
        }
      }
    }
    if (min_inner_rel > 0. && n_inner > 0) {
      const G mrel = G(min_inner_rel);
      auto& L0 = inner_[p][0];
      size_t ll = 0;
      while (ll < L0.size()) {
        const size_t l = L0[ll].cell;
        const int n = faceNeighbours(P, l, nb);
        G ming = var[nb[0]], maxg = var[nb[0]];
        for (int a = 1; a < n; ++a) {
          if (ming > var[nb[a]]) ming = var[nb[a]];
          if (maxg < var[nb[a]]) maxg = var[nb[a]];
        }
        const G gdiff = 0.5 * (maxg - ming);
        const G mind = mrel * gdiff;
        if (var[l] > -mind) {
          L0.erase(L0.begin() + ll);
          marker[l] = 0;
          for (int a = 0; a < n; ++a) {
            if (!marker[nb[a]] && var[nb[a]] < var[l]) {
              L0.push_back(mk(nb[a]));
              marker[nb[a]] = 1;
            }
          }
        } else {
          ++ll;
        }
      }
    }
    for (int layer = 1; layer < n_inner; ++layer)
      for (size_t ll = 0; ll < inner_[p][layer - 1].size(); ++ll) {
        const int n = faceNeighbours(P, inner_[p][layer - 1][ll].cell, nb);
        for (int a = 0; a < n; ++a)
          if (!marker[nb[a]] && var[nb[a]] < 0.) {
            inner_[p][layer].push_back(mk(nb[a]));
            marker[nb[a]] = 1;
          }
      }
    for (int layer = 1; layer < n_outer; ++layer)
      for (size_t ll = 0; ll < outer_[p][layer - 1].size(); ++ll) {
        const int n = faceNeighbours(P, outer_[p][layer - 1][ll].cell, nb);
        for (int a = 0; a < n; ++a)
          if (!marker[nb[a]] && var[nb[a]] > 0.) {
            outer_[p][layer].push_back(mk(nb[a]));
            marker[nb[a]] = 1;
          }
      }
    // normalised gradient (computeNablaVar, cartesianpatch.cpp:1592-1657) and mirror tables
    auto gradient = [&](size_t l, G* g) {
      size_t c[3];
      P.ijk(l, c[0], c[1], c[2]);
      const size_t n[3] = {(size_t)P.ni(), (size_t)P.nj(), (size_t)P.nk()};
      for (int a = 0; a < 3; ++a) {
        size_t cp[3] = {c[0], c[1], c[2]}, cm[3] = {c[0], c[1], c[2]};
        cp[a] = c[a] + 1;
        cm[a] = c[a] - 1;
        const bool skip_p = cp[a] > n[a] - 1, skip_m = cm[a] > n[a] - 1;
        const size_t lp = P.index(cp[0], cp[1], cp[2]), lm = P.index(cm[0], cm[1], cm[2]);
        if (!skip_p && !skip_m)
          g[a] = (var[lp] - var[lm]) / (2. * P.delta(a));
        else if (skip_p)
          g[a] = (var[l] - var[lm]) / P.delta(a);
        else
          g[a] = (var[lp] - var[l]) / P.delta(a);
      }
      G len = 0;
      for (int a = 0; a < 3; ++a) len += g[a] * g[a];
      len = std::sqrt(len);
      for (int a = 0; a < 3; ++a) g[a] /= len;
    };
    for (int layer = 0; layer < n_inner; ++layer) {
      std::vector<LayerCell>& L = inner_[p][layer];
      std::vector<LayerCell> kept;
      for (LayerCell& c : L) {
        G g[3];
        gradient(c.cell, g);
        c.gx = (double)g[0];
        c.gy = (double)g[1];
        c.gz = (double)g[2];
        const G shift_len = -2. * var[c.cell];
        G x, y, z;
        P.xyzCell(c.cell, x, y, z);
        WeightSet<G> ws;
        const int r = P.interpolCoeffs(x + shift_len * g[0], y + shift_len * g[1], z + shift_len * g[2], ws);
        if (r < 0) return -1;
        // tranferToFixedArrays(8) (weightedset.h:623-673); the reference pads with an
        // out-of-range entry, here the first donor with weight 0
        if (r == 1 && ws.n > 0) {
          for (int a = 0; a < 8; ++a) {
            c.donor[a] = a < ws.n ? ws.idx[a] : ws.idx[0];
            c.weight[a] = a < ws.n ? (double)ws.w[a] : 0.;
          }
          kept.push_back(c);
        }
      }
      L.swap(kept);
    }
    for (int layer = 0; layer < n_outer; ++layer)
      for (LayerCell& c : outer_[p][layer]) {
        G g[3];
        gradient(c.cell, g);
        c.gx = (double)g[0];
        c.gy = (double)g[1];
        c.gz = (double)g[2];
      }
  }
  return 0;
}

template struct WeightSet<float>;
template struct WeightSet<double>;
template struct Transform<float>;
template struct Transform<double>;
template void setupTransformation<float>(const float*, const float*, const float*, Transform<float>&,
                                         Transform<float>&);
template void setupTransformation<double>(const double*, const double*, const double*, Transform<double>&,
                                          Transform<double>&);
template class CartesianPatch<float>;
template class CartesianPatch<double>;
template class PatchGrid<float>;
template class PatchGrid<double>;

}  // namespace drnum
