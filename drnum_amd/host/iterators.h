// iterators.h -- C++ drop-in for the reference's GPU iterator path, running on
// libdrnhip through its C ABI (include/drnhip.h).
//
// The reference's path (iterators/patchiterator.h:28-72, tpatchiterator.h:29-61,
// gpu_cartesianiterator.h:29-304, rungekutta.cpp:27-42, timeintegration.cpp:28-55)
// is mirrored class by class with the same names and argument meaning:
//
//   PatchIterator            compute(factor, patches) / computeAll(factor) /
//                            copyField(src, dst) / copyDonorData(field) /
//                            updateHost() / updateDevice() / addPatch(patch)
//   HIP_CartesianIterator<DIM, OP>(OP op, int device = 0, size_t thread_limit = 0)
//                            the GPU_CartesianIterator<DIM, OP> replacement
//   IteratorFeeder           addIterator / feed: patch groups to the iterators
//                            of equal code string (iteratorfeeder.cpp:23-62)
//   GenericOperation         post operation (genericoperation.h:24-31)
//   HIP_LevelSetBC           level-set fill as a post operation (the
//                            CPU_LevelSetObjectBC<LSOBCCompressibleEulerOp> role)
//   TimeIntegration          addIterator / addPostOperation / copyField /
//                            computeIterators / copyDonorData / runPostOperations
//   RungeKutta               addAlpha, operator()(dt): the reference's stage loop
//   HIP_RungeKutta           the same step through drn_rk_step (no field copy,
//                            fewer exchanges) or a captured graph (drn_rk_graph_run)
//
// Patches are HostPatch<real> objects: the state arrays m_Data laid out
// [field][var][cell] as Patch::allocateData (patch.cpp:703-715), on top of a
// drnum::CartesianPatch<double> (geometry, seek layers, face functors) of a
// drnum::PatchGrid<double> that holds the donor tables.
//
// The flux operator OP is a type, as in the reference.  op_traits<OP> maps the
// supported operator types onto the library's operator enums at compile time;
// any other OP fails to compile.  Errors of the library raise drnum::DeviceError
// (the reference stops the process on BUG/ERROR, drnum.h:109-128; an uncaught
// DeviceError does the same).
#pragma once

#include <cstddef>
#include <cstdio>
#include <list>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/drnhip.h"
#include "../../include/drnum_host.h"
#include "grid.h"

namespace drnum {

// ---------------------------------------------------------------------------
// operator types (names of the reference's templates; no host arithmetic here)
// ---------------------------------------------------------------------------

struct VanAlbada {};     // reconstruction/vanalbada.h:28-33
struct MinMod {};        // reconstruction/minmod.h:28-31
struct SecondOrder {};   // reconstruction/secondorder.h:28-31
template <unsigned DIM, class TLimiter>
struct Upwind2 {};       // reconstruction/upwind2.h:29-52
template <unsigned DIM>
struct Upwind1 {};       // reconstruction/upwind1.h:29-39
struct PerfectGas {};    // perfectgas.h:31-34
template <unsigned DIM, class TReconstruction, class TGas>
struct AusmPlus {};      // fluxes/ausmplus.h:39-106
template <unsigned DIM, class TGas>
struct CompressibleViscFlux {};  // fluxes/compressibleviscflux.h:27-268
// An Euler flux followed by a viscous flux on every interior face, the
// composition the reference's applications write as their own flux class
// (EaFlux::xField/yField/zField, drnumCylinder/main.h:127-163).
template <class TEuler, class TViscous>
struct EulerViscFlux {};

template <class L>
struct limiter_code;
template <>
struct limiter_code<VanAlbada> {
  static constexpr int value = DRN_LIM_VANALBADA;
};
template <>
struct limiter_code<MinMod> {
  static constexpr int value = DRN_LIM_MINMOD;
};
template <>
struct limiter_code<SecondOrder> {
  static constexpr int value = DRN_LIM_SECONDORDER;
};

template <class OP, class Enable = void>
struct op_traits {
  static_assert(sizeof(OP) == 0, "this flux operator has no HIP implementation");
};
// An application's own flux class (the reference's JetFlux, EaFlux, ...) names
// the interior flux it composes as `interior_flux`, one of the types below.
template <class OP>
struct op_traits<OP, std::void_t<typename OP::interior_flux>> : op_traits<typename OP::interior_flux> {};
template <class L>
struct op_traits<AusmPlus<5, Upwind2<5, L>, PerfectGas>> {
  static constexpr int recon = DRN_RECON_UPWIND2;
  static constexpr int limiter = limiter_code<L>::value;
  static constexpr int viscous = 0;
};
template <>
struct op_traits<AusmPlus<5, Upwind1<5>, PerfectGas>> {
  static constexpr int recon = DRN_RECON_UPWIND1;
  static constexpr int limiter = DRN_LIM_VANALBADA;  // unused by first order
  static constexpr int viscous = 0;
};
template <class TEuler>
struct op_traits<EulerViscFlux<TEuler, CompressibleViscFlux<5, PerfectGas>>> {
  static constexpr int recon = op_traits<TEuler>::recon;
  static constexpr int limiter = op_traits<TEuler>::limiter;
  static constexpr int viscous = 1;
};

// The face functors of a flux class, per patch.  The reference's OP decides in
// its wall methods which functor a patch face gets (JetFlux::xWallM: the jet on
// the jet patch, a slip wall elsewhere, jetflux.h:158-192).  An OP that has
//   bool faces(const CartesianPatch<double>& geometry, size_t patch_index,
//              int32_t bc[6], double par[6][DRN_BC_NPAR]) const;
// sets the DRN_BC_* functor and parameters of every face of a patch there
// (returning true); the grid's face codes are used otherwise.  The far-field
// state is given in primitive variables p, T, u, v, w.
template <class OP, class = void>
struct has_faces : std::false_type {};
template <class OP>
struct has_faces<OP, std::void_t<decltype(std::declval<const OP&>().faces(
                         std::declval<const CartesianPatch<double>&>(), size_t(0), (int32_t*)nullptr,
                         (double(*)[DRN_BC_NPAR]) nullptr))>> : std::true_type {};

class DeviceError : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

// ---------------------------------------------------------------------------
// host patch: state arrays + geometry (Patch::m_Data, patch_common.h:26-95)
// ---------------------------------------------------------------------------

template <typename real>
class HostPatch {
 public:
  HostPatch(const PatchGrid<double>* grid, size_t index, size_t num_fields = 3, size_t num_vars = 5)
      : grid_(grid), index_(index), nf_(num_fields), nv_(num_vars) {
    data_.assign(nf_ * nv_ * grid_->patch(index_).cells(), real(0));
  }
  const CartesianPatch<double>& geometry() const { return grid_->patch(index_); }
  const PatchGrid<double>* grid() const { return grid_; }
  size_t getIndex() const { return index_; }
  size_t sizeI() const { return (size_t)geometry().ni(); }
  size_t sizeJ() const { return (size_t)geometry().nj(); }
  size_t sizeK() const { return (size_t)geometry().nk(); }
  size_t numFields() const { return nf_; }
  size_t numVariables() const { return nv_; }
  size_t variableSize() const { return geometry().cells(); }
  size_t fieldSize() const { return nv_ * variableSize(); }
  size_t dataSize() const { return nf_ * fieldSize(); }
  real* getData() { return data_.data(); }
  real* getField(size_t i_field) { return data_.data() + i_field * fieldSize(); }
  real* getVariable(size_t i_field, size_t i_var) { return getField(i_field) + i_var * variableSize(); }
  // Patch::setFieldToConst
  void setFieldToConst(size_t i_field, const real* var) {
    for (size_t v = 0; v < nv_; ++v) {
      real* p = getVariable(i_field, v);
      for (size_t l = 0; l < variableSize(); ++l) p[l] = var[v];
    }
  }

  // Per-cell flags, the patch's m_Active and m_IsInsideCell (patch_common.h:31,56):
  // an inactive cell keeps its state through the update (cartesianiterator.h:315),
  // and a face between an inside and an outside cell is a split face, where the
  // MUSCL reconstruction keeps first order (upwind2.h:41).  The iterator hands
  // them to the device with the state (updateDevice).
  void deactivate(size_t l) { flag(l) |= DRN_CELL_INACTIVE; }  // patch_common.h:505
  void activate(size_t l) {                                       // patch_common.h:506
    if (!flags_.empty()) flags_[l] &= (uint8_t)~DRN_CELL_INACTIVE;
  }
  bool isActive(size_t l) const { return flags_.empty() || !(flags_[l] & DRN_CELL_INACTIVE); }  // :507
  void deactivate(size_t i, size_t j, size_t k) { deactivate(geometry().index(i, j, k)); }  // cartesianpatch_common.h:332
  bool isInsideCell(size_t l) const { return !flags_.empty() && (flags_[l] & DRN_CELL_INSIDE); }  // patch_common.h:485
  bool isSplitFace(size_t a, size_t b) const { return isInsideCell(a) != isInsideCell(b); }     // patch_common.h:500
  // Patch::setSplitFaces (patch.h:895-931): the cells of the split faces are
  // split cells, those of faces marked `inside` are inside cells, and the
  // inside is then filled through the face neighbours (CartesianPatch::
  // getNeighbours, cartesianpatch.cpp:1841-1866) up to the split cells.  C is
  // a container of elements with `idx` and `inside` (splitface_t).
  template <class C>
  void setSplitFaces(const C& split_faces) {
    const CartesianPatch<double>& G = geometry();
    std::vector<char> split(variableSize(), 0);
    std::vector<size_t> front, next;
    for (const auto& f : split_faces) {
      split[f.idx] = 1;
      if (f.inside) {
        flag(f.idx) |= DRN_CELL_INSIDE;
        front.push_back(f.idx);
      }
    }
    while (!front.empty()) {
      next.clear();
      for (size_t l : front) {
        size_t i, j, k;
        G.ijk(l, i, j, k);
        const size_t nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
        for (const auto& q : nb) {
          if (!G.inRange(q[0], q[1], q[2])) continue;  // size_t wrap-around of -1 is out of range
          const size_t m = G.index(q[0], q[1], q[2]);
          if (!split[m] && !isInsideCell(m)) {
            flag(m) |= DRN_CELL_INSIDE;
            next.push_back(m);
          }
        }
      }
      front.swap(next);
    }
  }
  // the flags in the layout drn_set_cell_flags takes; null while every cell is
  // active and outside
  const uint8_t* cellFlags() const { return flags_.empty() ? nullptr : flags_.data(); }

 private:
  uint8_t& flag(size_t l) {
    if (flags_.empty()) flags_.assign(variableSize(), 0);
    return flags_[l];
  }
  const PatchGrid<double>* grid_;
  size_t index_, nf_, nv_;
  std::vector<real> data_;
  std::vector<uint8_t> flags_;
};

// splitface_t (splitface_t.h:24-38): the fields setSplitFaces reads are idx and inside
struct splitface_t {
  size_t idx = 0, idx_neigh = 0;
  double nx = 0, ny = 0, nz = 0, wnx = 0, wny = 0, wnz = 0, dist = 0, h1 = 0, h2 = 0;
  bool inside = false;
};

// PatchGrid::writeData / readData (patchgrid.cpp:563-610): field i_field of
// every patch to / from "<base>_<count, 6 digits>.dnd" (write) or the named
// file (read, ".dnd" appended when missing); readData returns the time.
template <typename real>
void writeData(std::vector<HostPatch<real>*>& patches, size_t i_field, const std::string& base_file_name, real time,
               int count) {
  char num[16];
  snprintf(num, sizeof(num), "%06d", count);
  const std::string name = base_file_name + "_" + num + ".dnd";
  std::vector<const void*> ptr;
  std::vector<int64_t> cnt;
  for (auto* P : patches) {
    ptr.push_back(P->getField(i_field));
    cnt.push_back((int64_t)P->fieldSize());
  }
  if (drnh_dnd_write(name.c_str(), (int)sizeof(real), (double)time, (int)patches.size(), ptr.data(), cnt.data()))
    throw DeviceError("writeData: cannot write " + name);
}

template <typename real>
real readData(std::vector<HostPatch<real>*>& patches, size_t i_field, const std::string& file_name) {
  std::vector<void*> ptr;
  std::vector<int64_t> cnt;
  for (auto* P : patches) {
    ptr.push_back(P->getField(i_field));
    cnt.push_back((int64_t)P->fieldSize());
  }
  double t = 0;
  int prec = 0;
  if (drnh_dnd_read(file_name.c_str(), (int)sizeof(real), &t, (int)patches.size(), ptr.data(), cnt.data(), &prec))
    throw DeviceError("readData: cannot read " + file_name + " for this grid");
  return (real)t;
}

// CompressibleVariables<TGas> (compressiblevariables.h:29-105): the
// post-processing quantities Ma, p, T, rho, S and the vector U, with the
// reference pressure and temperature of the entropy S (1e5 and 300 by default).
// The values themselves are computed on the device by drn_post_variables.
template <class TGas>
class CompressibleVariables {
 public:
  int numScalars() const { return 5; }
  int numVectors() const { return 1; }
  void setReferenceTemperature(double T) { ref_T_ = T; }
  void setReferencePressure(double p) { ref_p_ = p; }
  double referenceTemperature() const { return ref_T_; }
  double referencePressure() const { return ref_p_; }
  std::string getScalarName(int i) const {
    static const char* const names[5] = {"Ma", "p", "T", "rho", "S"};
    if (i < 0 || i >= 5) throw std::out_of_range("CompressibleVariables: scalar index");
    return names[i];
  }
  std::string getVectorName(int i) const {
    if (i != 0) throw std::out_of_range("CompressibleVariables: vector index");
    return "U";
  }

 private:
  double ref_p_ = 1e5, ref_T_ = 300;
};

// ---------------------------------------------------------------------------
// iterators
// ---------------------------------------------------------------------------

template <typename real>
class PatchIterator {
 public:
  virtual ~PatchIterator() {}
  virtual void updateHost() = 0;
  virtual void updateDevice() = 0;
  size_t numPatches() const { return patches_.size(); }
  HostPatch<real>* getPatch(size_t i) { return patches_[i]; }
  // PatchIterator::computeAll (patchiterator.h:80-88): every active patch
  void computeAll(real factor) {
    std::vector<size_t> ids;
    for (size_t i = 0; i < patches_.size(); ++i)
      if (active_[i]) ids.push_back(i);
    compute(factor, ids);
  }
  virtual void addPatch(HostPatch<real>* patch) {
    patches_.push_back(patch);
    active_.push_back(true);
  }
  virtual void compute(real factor, const std::vector<size_t>& patches) = 0;
  virtual void copyField(size_t i_src, size_t i_dst) = 0;
  virtual void copyDonorData(size_t i_field) = 0;
  void activatePatch(size_t i) { active_[i] = true; }
  void deactivatePatch(size_t i) { active_[i] = false; }
  bool patchActive(size_t i) const { return active_[i]; }
  // setCodeString / getCodeString (patchiterator.h:61-70, 96-104): the string
  // is normalised as a CodeString; the default is "void" (patchiterator.h:77)
  void setCodeString(const CodeString& s) { codes_ = s; }
  CodeString getCodeString() const { return codes_; }

 protected:
  std::vector<HostPatch<real>*> patches_;
  std::vector<bool> active_;
  CodeString codes_{"void"};
};

// IteratorFeeder (iteratorfeeder.cpp:23-62).  feed() walks the grid's patch
// groups; every iterator whose code string equals a group's receives all of
// the group's patches, in group order.  A group no iterator matches stops the
// run (the reference prints the code string and calls BUG): here FeedError,
// naming the code string.  The reference's patches live in its PatchGrid; here
// the caller passes the HostPatch of every grid patch, indexed as the grid.
class FeedError : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

template <typename real>
class IteratorFeeder {
 public:
  void addIterator(PatchIterator<real>* iterator) { iterators_.push_back(iterator); }
  void feed(const PatchGrid<double>& grid, const std::vector<HostPatch<real>*>& patches) {
    if (patches.size() != grid.numPatches()) throw FeedError("feed: one HostPatch per grid patch expected");
    const PatchGroups& groups = grid.getPatchGroups();
    for (size_t g = 0; g < groups.accessNumPatchGroups(); ++g) {
      const SinglePatchGroup& spg = groups.accessSinglePatchGroup(g);
      bool found = false;
      for (auto* it : iterators_) {
        if (it->getCodeString() != spg.solver_codes) continue;
        found = true;
        for (size_t p : spg.patches) it->addPatch(patches[p]);
      }
      if (!found) throw FeedError("no iterator for the code string \"" + spg.solver_codes + "\"");
    }
  }

 private:
  std::vector<PatchIterator<real>*> iterators_;
};

// GenericOperation (genericoperation.h:24-31): a post operation of the time integration
class GenericOperation {
 public:
  virtual ~GenericOperation() {}
  virtual void operator()() = 0;
};

template <typename real>
struct real_precision;
template <>
struct real_precision<float> {
  static constexpr int value = DRN_F32;
};
template <>
struct real_precision<double> {
  static constexpr int value = DRN_F64;
};

// GPU_CartesianIterator<DIM, OP> (gpu_cartesianiterator.h:29-304) on libdrnhip.
// One device context holds every patch added before the first device call; a
// stage is one launch over all of them.  As GPU_CartesianIterator::compute
// (gpu_cartesianiterator.h:224), compute() advances every patch whatever the
// `patches` argument holds.  thread_limit is accepted for interface parity
// and unused (the work list, not a block size, is tuned: drn_autotune).
template <unsigned DIM, class OP, typename real = float>
class HIP_CartesianIterator : public PatchIterator<real> {
  static_assert(DIM == 5, "the HIP path covers the 5-variable compressible system");

 public:
  explicit HIP_CartesianIterator(OP op = OP(), int device = 0, size_t thread_limit = 0) : op_(op) {
    (void)thread_limit;
    check(drn_ctx_create(device, real_precision<real>::value, &ctx_), "drn_ctx_create");
    check(drn_set_operator(ctx_, op_traits<OP>::recon, op_traits<OP>::limiter), "drn_set_operator");
    check(drn_set_viscous(ctx_, op_traits<OP>::viscous), "drn_set_viscous");
  }
  ~HIP_CartesianIterator() override { drn_ctx_destroy(ctx_); }
  HIP_CartesianIterator(const HIP_CartesianIterator&) = delete;
  HIP_CartesianIterator& operator=(const HIP_CartesianIterator&) = delete;

  drn_ctx* context() {
    setup();
    return ctx_;
  }

  void addPatch(HostPatch<real>* patch) override {
    if (laid_out_) throw DeviceError("HIP_CartesianIterator::addPatch after the first device call");
    PatchIterator<real>::addPatch(patch);
  }

  // Patch -> device copy of all fields (GPU_Patch::copyToDevice, gpu_patch.h:130-134)
  // and the per-cell flags (m_Active, m_IsInsideCell) of every patch
  void updateDevice() override {
    setup();
    for (size_t p = 0; p < this->patches_.size(); ++p) {
      for (size_t f = 0; f < this->patches_[p]->numFields() && f < 3; ++f)
        check(drn_patch_upload(ctx_, (int)p, (int)f, this->patches_[p]->getField(f)), "drn_patch_upload");
      const uint8_t* fl = this->patches_[p]->cellFlags();
      if (fl || flags_sent_) check(drn_set_cell_flags(ctx_, (int)p, fl), "drn_set_cell_flags");
      if (fl) flags_sent_ = true;
    }
  }
  // device -> Patch copy of all fields (GPU_Patch::copyFromDevice, gpu_patch.h:136-139)
  void updateHost() override {
    setup();
    for (size_t p = 0; p < this->patches_.size(); ++p)
      for (size_t f = 0; f < this->patches_[p]->numFields() && f < 3; ++f)
        check(drn_patch_download(ctx_, (int)p, (int)f, this->patches_[p]->getField(f)), "drn_patch_download");
  }

  // CartesianIterator::compute: field 0 <- field 1 + factor/dV * Res(field 0).  The
  // kernel writes the scratch field 2, which then becomes logical field 0
  // (field 2 is the reference's scratch / abuse field).
  void compute(real factor, const std::vector<size_t>&) override {
    setup();
    check(drn_stage(ctx_, 0, 1, 2, (double)factor, launch_), "drn_stage");
    check(drn_swap_fields(ctx_, 0, 2), "drn_swap_fields");
  }
  void copyField(size_t i_src, size_t i_dst) override {
    setup();
    check(drn_copy_field(ctx_, (int)i_src, (int)i_dst), "drn_copy_field");
  }
  void copyDonorData(size_t i_field) override {
    setup();
    check(drn_exchange(ctx_, (int)i_field), "drn_exchange");
  }

  // DRN_LAUNCH_BATCHED (default: one launch over all patches) or DRN_LAUNCH_PER_PATCH
  void setLaunchForm(int launch) { launch_ = launch; }
  int autotune() {
    setup();
    int choice = -1;
    check(drn_autotune(ctx_, launch_, &choice), "drn_autotune");
    return choice;
  }
  void sync() { check(drn_sync(ctx_), "drn_sync"); }
  // CFL_max of field i_field for time step dt as the jet demo's diagnostics
  // loop computes it (applications/drnumJetDemo/main.h:196-205): the maximum
  // over cells of max(|u|, |u+a|, |u-a|) dt/dx, x axis only, in `real`
  double cflMax(size_t i_field, double dt) {
    setup();
    double out[4] = {0, 0, 0, 0};
    check(drn_diagnostics(ctx_, (int)i_field, dt, out), "drn_diagnostics");
    return out[0];
  }
  // not in the reference: the CFL number over all three axes, max over cells
  // and axes of (|u_d| + a) dt / delta_d
  double cflAxes(size_t i_field, double dt) {
    setup();
    double c = 0;
    check(drn_cfl_axes(ctx_, (int)i_field, dt, &c), "drn_cfl_axes");
    return c;
  }
  // Patch::computeVariableDifference (patch.cpp:733-746) for the iterator's
  // patch i_patch, on the device: max_norm as the reference computes it, the
  // sum of squares accumulated in FP64
  void computeVariableDifference(size_t i_patch, size_t i_field1, size_t i_var1, size_t i_field2, size_t i_var2,
                                 real& max_norm, real& l2_norm) {
    setup();
    double mx = 0, l2 = 0;
    check(drn_variable_difference(ctx_, (int)i_patch, (int)i_field1, (int)i_var1, (int)i_field2, (int)i_var2, &mx, &l2),
          "drn_variable_difference");
    max_norm = (real)mx;
    l2_norm = (real)l2;
  }
  int launchForm() const { return launch_; }
  // the level-set fill registered on this iterator's context (HIP_LevelSetBC),
  // which drn_rk_step runs as its post operation
  const GenericOperation* levelSetOperation() const { return ls_op_; }
  void setLevelSetOperation(const GenericOperation* op) {
    if (ls_op_ && ls_op_ != op) throw DeviceError("HIP_CartesianIterator: a second level-set operation");
    ls_op_ = op;
  }

  // PatchGrid::writeToVtk (patchgrid.cpp:673-695): the donor data of field 0 are
  // exchanged first (accessAllDonorData(0)), then the core cells of every patch
  // of this iterator go to <file_name>[_<count, 6 digits>].vtm as block p (the
  // iterator's patch order), with the quantities of `vars` evaluated on the
  // device from field i_field (drn_post_variables)
  template <class TGas>
  void writeToVtk(size_t i_field, const std::string& file_name, const CompressibleVariables<TGas>& vars,
                  int count = -1) {
    setup();
    check(drn_exchange(ctx_, 0), "drn_exchange");
    std::vector<char> dir(file_name.size() + 64);
    if (drnh_vtk_write_multiblock(file_name.c_str(), count, (int)this->patches_.size(), dir.data(),
                                  (int)dir.size()) < 0)
      throw DeviceError("writeToVtk: cannot write " + file_name + ".vtm");
    const std::string base(dir.data());
    const size_t cut = base.find_last_of('/');
    const std::string leaf = cut == std::string::npos ? base : base.substr(cut + 1);
    std::vector<std::string> names;
    for (int i = 0; i < vars.numScalars(); ++i) names.push_back(vars.getScalarName(i));
    const std::string vname = vars.getVectorName(0);
    for (size_t p = 0; p < this->patches_.size(); ++p) {
      const HostPatch<real>& hp = *this->patches_[p];
      const CartesianPatch<double>& P = hp.geometry();
      const size_t N = hp.variableSize();
      std::vector<real> q(8 * N);
      check(drn_post_variables(ctx_, (int)i_field, (int)p, 0xFF, vars.referencePressure(),
                               vars.referenceTemperature(), q.data()),
            "drn_post_variables");
      const std::vector<float> qf(q.begin(), q.end());
      const float* sc[5];
      const char* sn[5];
      for (int i = 0; i < 5; ++i) {
        sc[i] = qf.data() + (size_t)i * N;
        sn[i] = names[(size_t)i].c_str();
      }
      const float* vec[1] = {qf.data() + 5 * N};
      const char* vn[1] = {vname.c_str()};
      const int32_t n3[3] = {P.ni(), P.nj(), P.nk()};
      int32_t seek6[6];
      double o[3], d[3];
      for (int f = 0; f < 6; ++f) seek6[f] = P.seek(f);
      for (int a = 0; a < 3; ++a) {
        o[a] = P.origin(a);
        d[a] = P.delta(a);
      }
      const std::string path = base + "/" + leaf + "_" + std::to_string(p) + ".vts";
      if (drnh_vtk_write_patch(path.c_str(), n3, seek6, o, d, 5, sn, sc, 1, vn, vec))
        throw DeviceError("writeToVtk: cannot write " + path);
    }
  }
  const OP& op() const { return op_; }

  void check(int rc, const char* what) const {
    if (rc != DRN_OK) throw DeviceError(std::string(what) + ": " + (ctx_ ? drn_last_error(ctx_) : "no context"));
  }

 private:
  // device layout of every added patch: geometry, face functors (farfield state
  // converted to conserved variables in `real`), donor tables of the grid
  void setup() {
    if (laid_out_) return;
    if (this->patches_.empty()) throw DeviceError("HIP_CartesianIterator: no patches");
    const PatchGrid<double>* grid = this->patches_[0]->grid();
    for (size_t p = 0; p < this->patches_.size(); ++p) {
      const HostPatch<real>& hp = *this->patches_[p];
      if (hp.grid() != grid) throw DeviceError("HIP_CartesianIterator: patches of different grids");
      const CartesianPatch<double>& P = hp.geometry();
      drn_patch_desc d{};
      d.ni = P.ni();
      d.nj = P.nj();
      d.nk = P.nk();
      for (int f = 0; f < 6; ++f) d.seek[f] = P.seek(f);
      // CartesianPatch::computeDeltas in `real` (cartesianpatch.cpp:164-168)
      d.dx = (double)(real(P.length(0)) / real((size_t)P.ni()));
      d.dy = (double)(real(P.length(1)) / real((size_t)P.nj()));
      d.dz = (double)(real(P.length(2)) / real((size_t)P.nk()));
      for (int f = 0; f < 6; ++f) {
        d.bc[f] = P.face[f];
        for (int q = 0; q < DRN_BC_NPAR; ++q) d.bc_par[f][q] = P.face_par[f][q];
      }
      if constexpr (has_faces<OP>::value) {
        int32_t bc[6];
        double par[6][DRN_BC_NPAR];
        for (int f = 0; f < 6; ++f) {
          bc[f] = d.bc[f];
          for (int q = 0; q < DRN_BC_NPAR; ++q) par[f][q] = d.bc_par[f][q];
        }
        if (op_.faces(P, hp.getIndex(), bc, par)) {
          for (int f = 0; f < 6; ++f) {
            d.bc[f] = bc[f];
            for (int q = 0; q < DRN_BC_NPAR; ++q) d.bc_par[f][q] = par[f][q];
          }
        }
      }
      for (int f = 0; f < 6; ++f) {
        if (d.bc[f] == FACE_FARFIELD) {
          // PerfectGas::primitiveToConservative in `real` (perfectgas.h:56-64)
          const real pr = real(d.bc_par[f][0]), T = real(d.bc_par[f][1]), u = real(d.bc_par[f][2]),
                     v = real(d.bc_par[f][3]), w = real(d.bc_par[f][4]);
          const real rho = pr / (real(287.2) * T);
          d.bc_par[f][0] = (double)rho;
          d.bc_par[f][1] = (double)(real)(rho * u);
          d.bc_par[f][2] = (double)(real)(rho * v);
          d.bc_par[f][3] = (double)(real)(rho * w);
          d.bc_par[f][4] = (double)(real)(rho * (real(717.5) * T + 0.5 * (u * u + v * v + w * w)));
        }
      }
      int id = -1;
      check(drn_patch_add(ctx_, &d, &id), "drn_patch_add");
    }
    check(drn_layout(ctx_), "drn_layout");
    // donor tables (Patch::buildDonorTransferData, patch.cpp:128-218), for patches of this iterator
    bool any = false;
    for (size_t p = 0; p < this->patches_.size(); ++p) {
      const DonorTables& t = grid->tables(this->patches_[p]->getIndex());
      if (t.donors.empty()) continue;
      std::vector<int32_t> dp;
      std::vector<int64_t> info, ru(t.receivers.begin(), t.receivers.end()),
          rc(t.rec_concat.begin(), t.rec_concat.end()), di(t.donor_index.begin(), t.donor_index.end());
      for (const auto& D : t.donors) {
        int local = -1;
        for (size_t q = 0; q < this->patches_.size(); ++q)
          if ((int)this->patches_[q]->getIndex() == D.patch) local = (int)q;
        if (local < 0) throw DeviceError("HIP_CartesianIterator: a donor patch is not in this iterator");
        dp.push_back(local);
        info.push_back((int64_t)D.nrec);
        info.push_back((int64_t)D.stride);
        info.push_back((int64_t)D.rec_start);
        info.push_back((int64_t)D.wi_start);
      }
      check(drn_exchange_set(ctx_, (int)p, (int)t.donors.size(), dp.data(), info.data(), (int64_t)ru.size(),
                             ru.data(), (int64_t)rc.size(), rc.data(), (int64_t)di.size(), di.data(),
                             t.donor_weight.data()),
            "drn_exchange_set");
      // the momentum's turn of every donor (donor_t axx..azz, patch.cpp:178-186, 627-641)
      std::vector<double> turn;
      for (const auto& D : t.donors) turn.insert(turn.end(), D.turn, D.turn + 9);
      check(drn_exchange_set_turn(ctx_, (int)p, (int)t.donors.size(), turn.data()), "drn_exchange_set_turn");
      any = true;
    }
    if (any) check(drn_exchange_commit(ctx_), "drn_exchange_commit");
    laid_out_ = true;
  }

  OP op_;
  drn_ctx* ctx_ = nullptr;
  bool laid_out_ = false;
  bool flags_sent_ = false;  // some patch has had flags on the device
  int launch_ = DRN_LAUNCH_BATCHED;
  const GenericOperation* ls_op_ = nullptr;
};

// ---------------------------------------------------------------------------
// post operations and time integration
// ---------------------------------------------------------------------------

// The OP of CPU_LevelSetObjectBC<DIM, OP>: the slip wall
// (lsobccompressibleeulerop.h:49-144) or the simple wall
// (lsobccompressibleswallop.h:48-116), as type tags
struct LSOBCCompressibleEulerOp {};
struct LSOBCCompressibleSWallOp {};
template <class L>
struct ls_op_code;
template <>
struct ls_op_code<LSOBCCompressibleEulerOp> {
  static constexpr int value = DRN_LS_SLIP;
};
template <>
struct ls_op_code<LSOBCCompressibleSWallOp> {
  static constexpr int value = DRN_LS_SWALL;
};

// The level-set fill (CPU_LevelSetObjectBC<LSOP>, cpu_levelsetobjectbc.h:58-90,
// relax 0.5) on the inner layers the grid's level-set builder produced, for the
// patches of one HIP iterator; LSOP is one of the tags above.
template <unsigned DIM, class OP, typename real, class LSOP = LSOBCCompressibleEulerOp>
class HIP_LevelSetBC : public GenericOperation {
 public:
  HIP_LevelSetBC(HIP_CartesianIterator<DIM, OP, real>* it, size_t field = 0, double relax = 0.5)
      : it_(it), field_(field), relax_(relax) {
    drn_ctx* ctx = it_->context();
    const PatchGrid<double>* grid = it_->getPatch(0)->grid();
    for (size_t p = 0; p < it_->numPatches(); ++p) {
      std::vector<int64_t> cells, donors;
      std::vector<double> nrm, w;
      const size_t gp = it_->getPatch(p)->getIndex();
      for (int layer = 0; layer < grid->numInnerLayers(); ++layer)
        for (const LayerCell& c : grid->innerLayer(gp, layer)) {
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
      it_->check(drn_levelset_set(ctx, (int)p, (int64_t)cells.size(), cells.data(), nrm.data(), donors.data(),
                                  w.data()),
                 "drn_levelset_set");
    }
    it_->check(drn_levelset_commit(ctx), "drn_levelset_commit");
    it_->check(drn_set_levelset_op(ctx, ls_op_code<LSOP>::value), "drn_set_levelset_op");
    it_->setLevelSetOperation(this);
  }
  void operator()() override { it_->check(drn_levelset_fill(it_->context(), (int)field_, relax_), "drn_levelset_fill"); }

 private:
  HIP_CartesianIterator<DIM, OP, real>* it_;
  size_t field_;
  double relax_;
};

// TimeIntegration (timeintegration.cpp:28-55)
template <typename real>
class TimeIntegration {
 public:
  virtual ~TimeIntegration() {}
  void addIterator(PatchIterator<real>* it) { iterators_.push_back(it); }
  void addPostOperation(GenericOperation* op) { post_.push_back(op); }
  void copyField(size_t i_src, size_t i_dst) {
    for (auto* it : iterators_) it->copyField(i_src, i_dst);
  }
  void computeIterators(real factor) {
    for (auto* it : iterators_) it->computeAll(factor);
  }
  void copyDonorData(size_t i_field) {
    for (auto* it : iterators_) it->copyDonorData(i_field);
  }
  void runPostOperations() {
    for (auto* op : post_) (*op)();
  }
  virtual void operator()(real dt) = 0;

 protected:
  std::list<PatchIterator<real>*> iterators_;
  std::list<GenericOperation*> post_;
};

// RungeKutta (rungekutta.cpp:27-42): the reference's stage loop over the virtual
// iterator interface (copy of field 0 into field 1, two exchanges per stage).
template <typename real>
class RungeKutta : public TimeIntegration<real> {
 public:
  void addAlpha(real a) { alpha_.push_back(a); }
  void operator()(real dt) override {
    bool first_step = true;
    this->copyDonorData(0);
    this->copyField(0, 1);
    for (real a : alpha_) {
      if (!first_step) this->copyDonorData(0);
      first_step = false;
      this->computeIterators(a * dt);
      this->copyDonorData(0);
      this->runPostOperations();
    }
  }
  const std::list<real>& alphas() const { return alpha_; }

 protected:
  std::list<real> alpha_;
};

// The same step for one HIP iterator in a single library call: the stages
// alternate between the two spare state buffers instead of copying field 0 into
// field 1, and the exchange before a stage is skipped when nothing changed the
// field since the last one.  A level-set fill set on the iterator's context runs
// after every stage's exchange, as the post operation does.  graph(dt, n) runs
// n steps from captured HIP graphs.
//
// The library call runs exactly the one iterator and the level-set fill
// registered on its context, so the step refuses (DeviceError) any other
// iterator or post operation added with addIterator / addPostOperation, and a
// registered fill that was not added as a post operation: run such a setup
// with RungeKutta, the reference's stage loop over the virtual interface.
template <unsigned DIM, class OP, typename real>
class HIP_RungeKutta : public RungeKutta<real> {
 public:
  explicit HIP_RungeKutta(HIP_CartesianIterator<DIM, OP, real>* it) : it_(it) { this->addIterator(it); }
  void operator()(real dt) override {
    checkSetup();
    std::vector<double> a(this->alpha_.begin(), this->alpha_.end());
    it_->check(drn_rk_step(it_->context(), (double)dt, (int)a.size(), a.data(), it_->launchForm()), "drn_rk_step");
  }
  void graph(real dt, int n_steps) {
    checkSetup();
    std::vector<double> a(this->alpha_.begin(), this->alpha_.end());
    it_->check(drn_rk_graph_run(it_->context(), (double)dt, (int)a.size(), a.data(), n_steps), "drn_rk_graph_run");
  }

 private:
  void checkSetup() const {
    if (this->iterators_.size() != 1 || this->iterators_.front() != it_)
      throw DeviceError("HIP_RungeKutta runs its one HIP iterator only; use RungeKutta for several iterators");
    const GenericOperation* ls = it_->levelSetOperation();
    size_t n_ls = 0;
    for (const GenericOperation* op : this->post_) {
      if (op != ls || ls == nullptr)
        throw DeviceError("HIP_RungeKutta runs no post operation but the iterator's level-set fill; use RungeKutta");
      ++n_ls;
    }
    if (ls != nullptr && n_ls != 1)
      throw DeviceError("HIP_RungeKutta: the iterator's level-set fill must be added once with addPostOperation");
  }

  HIP_CartesianIterator<DIM, OP, real>* it_;
};

}  // namespace drnum
