// grid.h -- host-side patch grid of the drop-in: Cartesian patch geometry,
// donor (inter-patch interpolation) tables and level-set layer tables.
//
// Geometry is templated on the coordinate type G.  The product uses G = double
// (FP64 geometry; weights are rounded to the state precision when handed to
// the device).  G = float reproduces the reference's default build, whose
// `real` is float (drnum.h:43), and is used to check the builder against it.
//
// Patches carry their base vectors (patch.cpp:63-85); the relative transform
// between two patches and the turn of the momentum are composed exactly as the
// reference's CoordTransform / CoordTransformVV compose them
// (math/coordtransform.{h,cpp}, math/coordtransformvv.cpp).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "codestring.h"

namespace drnum {

// face functor codes (include/drnhip.h DRN_BC_*)
enum FaceCode { FACE_SLIP = 0, FACE_FARFIELD = 1, FACE_TRANSMISSIVE = 2, FACE_JET = 3, FACE_NONE = 4 };

// A small weighted set of (cell index, weight) pairs: the subset of
// WeightedSet<T> (weightedset.h) used by the interpolation coefficients.
template <typename G>
struct WeightSet {
  int n = 0;
  size_t idx[8];
  G w[8];
  void unify();                      // weightedset.h:374-404
  void eliminateBelow(G eps);        // weightedset.h:569-602 (absolute, no sum keeping)
  void shiftToSum(G sum);            // weightedset.h:605-619
};

// CoordTransform (math/coordtransform.h:26-94): x -> A x + b
template <typename G>
struct Transform {
  G a[3][3], b[3];
  static Transform identity();                             // coordtransform.cpp:23-28
  void setBaseIJ(const G* base_i, const G* base_j);        // setMatrixFromBaseIJ, coordtransform.cpp:37-58
  Transform inverse() const;                               // coordtransform.cpp:99-114
  Transform concatenate(const Transform& c) const;         // coordtransform.h:136-165: this(c(x))
  void transform(G x, G y, G z, G* out) const;             // coordtransform.h:105-111
};

// Patch::setupTransformation (patch.cpp:265-278): the inertial -> this transform
// (m_TransformInertial2This) and its stored inverse (CoordTransformVV::computeInverse)
template <typename G>
void setupTransformation(const G* origin, const G* base_i, const G* base_j, Transform<G>& i2t, Transform<G>& t2i);

template <typename G>
class CartesianPatch {
 public:
  // base_i / base_j: base vectors of the local i and j axes in the inertial
  // frame; null means axis aligned (1,0,0) / (0,1,0)
  CartesianPatch(int ni, int nj, int nk, const int* seek, const double* origin, const double* length,
                 const double* base_i = nullptr, const double* base_j = nullptr);

  int ni() const { return n_[0]; }
  int nj() const { return n_[1]; }
  int nk() const { return n_[2]; }
  size_t cells() const { return (size_t)n_[0] * n_[1] * n_[2]; }
  size_t index(size_t i, size_t j, size_t k) const { return i * njk_ + j * (size_t)n_[2] + k; }
  void ijk(size_t l, size_t& i, size_t& j, size_t& k) const {
    i = l / njk_;
    size_t r = l - i * njk_;
    j = r / (size_t)n_[2];
    k = r - j * (size_t)n_[2];
  }
  bool inRange(size_t i, size_t j, size_t k) const {
    return i < (size_t)n_[0] && j < (size_t)n_[1] && k < (size_t)n_[2];
  }
  G delta(int a) const { return d_[a]; }
  G origin(int a) const { return o_[a]; }
  G baseI(int a) const { return bi_[a]; }
  G baseJ(int a) const { return bj_[a]; }
  const G* baseI() const { return bi_; }
  const G* baseJ() const { return bj_; }
  // m_TransformInertial2This and its stored inverse (this -> inertial, Patch::xyzoCell)
  const Transform<G>& inertial2this() const { return i2t_; }
  const Transform<G>& this2inertial() const { return t2i_; }
  // new base vectors (the orientation a grid file gives, patch.cpp:63-85)
  void setBase(const double* base_i, const double* base_j);
  G length(int a) const { return L_[a]; }
  int seek(int f) const { return seek_[f]; }
  // cell centre in patch coordinates (cartesianpatch.h:267-273)
  void xyzCell(size_t l, G& x, G& y, G& z) const;
  // computeCCDataInterpolCoeffs_V1 (cartesianpatch.cpp:1166-1299); returns 1 inside,
  // 0 outside, -1 on the reference's BUG conditions
  int interpolCoeffs(G x, G y, G z, WeightSet<G>& ws) const;
  // seek cells, sorted and unique (cartesianpatch.cpp:393-491, patch.cpp:468-486)
  std::vector<size_t> seekCells() const;
  bool isCore(size_t i, size_t j, size_t k) const;
  // false when no point of the box [lo, hi] (patch coordinates) can fall into
  // the donor zone, with a margin far above rounding: a candidate-pair prune
  // that changes no table (points outside the zone get no coefficients)
  bool donorZoneMayContain(const double* lo, const double* hi) const;
  // cell-centre bounding box of the whole patch in its own coordinates
  void cellCentreBox(double* lo, double* hi) const;

  // face functors and their parameters (drn_patch_desc)
  int face[6] = {FACE_SLIP, FACE_SLIP, FACE_SLIP, FACE_SLIP, FACE_SLIP, FACE_SLIP};
  double face_par[6][16] = {};
  std::string comment;
  CodeString solver_codes;  // Patch::m_solvercodes, read by Patch::readSolverCodes (patch.cpp:259-262)

 private:
  int lin(G s, int axis, size_t& ref, G& coeff) const;  // linearCCInterpolCoeffs
  int n_[3];
  size_t njk_;
  int seek_[6];
  G o_[3], L_[3], d_[3], ccmin_[3];
  G bi_[3], bj_[3];
  Transform<G> i2t_, t2i_;
  size_t dz_first_[3], dz_after_[3];
};

// Direct transfer lists of one receiver patch (patch.cpp:128-218, donor_t.h:28-45).
struct DonorTables {
  struct Donor {
    int patch;
    size_t nrec, stride, rec_start, wi_start;
    double turn[9];  // donor_t axx..azz (row-major): the momentum's turn donor -> receiver
  };
  std::vector<size_t> receivers;   // unique receiver cells (m_ReceivingCellIndicesUnique)
  std::vector<Donor> donors;
  std::vector<size_t> rec_concat;  // m_ReceivingCellIndicesConcat
  std::vector<size_t> donor_index; // m_DonorIndexConcat
  std::vector<double> donor_weight;// m_DonorWeightConcat (geometry precision, widened)
  size_t no_donor = 0;             // seek cells without any donor (warned, left alone)
};

// Level-set layer cell (LSLayerDataExtrapol, lslayerdataextrapol.h)
struct LayerCell {
  size_t cell;
  double g, gx, gy, gz;
  size_t donor[8];
  double weight[8];
};

template <typename G>
class PatchGrid {
 public:
  int addPatch(const CartesianPatch<G>& p);
  // PatchGrid::readGrid (patchgrid.cpp:479-561): "1001 {...}" entries, "0" ends;
  // every patch read also enters the patch groups (patchgrid.cpp:549)
  int readGrid(const std::string& path, double scale, std::string* err);
  // PatchGrid::getPatchGroups (patchgrid.h:310); groups allow solver exceptions
  // as the reference's grid constructs them (patchgrid.cpp:35)
  const PatchGroups& getPatchGroups() const { return groups_; }
  size_t numPatches() const { return patches_.size(); }
  CartesianPatch<G>& patch(size_t p) { return patches_[p]; }
  const CartesianPatch<G>& patch(size_t p) const { return patches_[p]; }

  // PatchGrid::computeDependencies (patchgrid.cpp:119-259).  period[a] > 0 adds
  // periodic images along axis a (an extension: the reference has no periodic
  // donor relations).  Returns 0, or -1 on a BUG condition of the reference.
  int computeDependencies(const double* period, int threads);
  const DonorTables& tables(size_t p) const { return tables_[p]; }

  // LevelSetObject::update + extractBCellLayers (levelsetobject.cpp:44-388)
  // for the level set min_s(|x - c_s| - r_s) (spherelevelset.cpp, combilevelsetor.cpp).
  int levelSetSpheres(const std::vector<double>& spheres, int n_inner, int n_outer, double min_inner_rel);
  const std::vector<G>& levelSetValues(size_t p) const { return lsg_[p]; }
  const std::vector<LayerCell>& innerLayer(size_t p, int layer) const { return inner_[p][layer]; }
  const std::vector<LayerCell>& outerLayer(size_t p, int layer) const { return outer_[p][layer]; }
  int numInnerLayers() const { return n_inner_; }
  int numOuterLayers() const { return n_outer_; }

  // PatchGrid::computeMinChLength (patchgrid.cpp:659-671)
  double minChLength() const;

 private:
  std::vector<CartesianPatch<G>> patches_;
  PatchGroups groups_{true};
  std::vector<DonorTables> tables_;
  std::vector<std::vector<G>> lsg_;
  std::vector<std::vector<std::vector<LayerCell>>> inner_, outer_;
  int n_inner_ = 0, n_outer_ = 0;
};

}  // namespace drnum
