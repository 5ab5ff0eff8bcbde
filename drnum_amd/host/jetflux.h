// jetflux.h -- the jet demo's flux operator (applications/drnumJetDemo/jetflux.h)
// for the HIP drop-in.
//
// JetFlux composes an interior flux with the face functors of the jet demo and
// chooses them per patch in its wall methods:
//   xWallM  the top-hat jet on the jet patch (jetflux.h:158-192), the slip wall
//           on every other patch;
//   xWallP  the slip wall (m_WallFlux, jetflux.h:140-144);
//   y and z walls the far-field functor (m_FarfieldFlux, jetflux.h:146-156,
//           193-198) with the state it is constructed with, 1e5 Pa and 300 K
//           at rest (compressiblefarfieldflux.h:75-78).
// Here those choices become DRN_BC_* codes per patch face (faces()), which
// HIP_CartesianIterator hands to the device.  The jet's centre is the middle
// of the patch's x-minus face, (0.5 sizeJ dy, 0.5 sizeK dz) as jetflux.h:170-171
// computes it, and it carries no velocity noise.  The interior flux is a type
// the library runs, AUSM+ with or without CompressibleViscFlux; the reference
// JetFlux's VanLeer flux is not on this path.
#pragma once

#include "iterators.h"

namespace drnum {

template <class TInterior = EulerViscFlux<AusmPlus<5, Upwind2<5, VanAlbada>, PerfectGas>,
                                          CompressibleViscFlux<5, PerfectGas>>>
class JetFlux {
 public:
  typedef TInterior interior_flux;

  // JetFlux::setup (jetflux.h:86-97)
  void setup(double u_jet, double u_far, double p_jet, double p_far, double T_jet, double T_far, double radius,
             int jet_patch, bool super_sonic) {
    u_jet_ = u_jet;
    u_far_ = u_far;
    p_jet_ = p_jet;
    p_far_ = p_far;
    T_jet_ = T_jet;
    T_far_ = T_far;
    radius_ = radius;
    jet_patch_ = jet_patch;
    super_sonic_ = super_sonic;
  }

  bool faces(const CartesianPatch<double>& P, size_t patch_index, int32_t bc[6], double par[6][DRN_BC_NPAR]) const {
    for (int f = 0; f < 6; ++f)
      for (int q = 0; q < DRN_BC_NPAR; ++q) par[f][q] = 0.0;
    bc[1] = DRN_BC_SLIP;
    if ((int)patch_index == jet_patch_) {
      bc[0] = DRN_BC_JET;
      par[0][0] = 0.5 * P.nj() * P.delta(1);  // y0 (jetflux.h:170)
      par[0][1] = 0.5 * P.nk() * P.delta(2);  // z0 (jetflux.h:171)
      par[0][2] = radius_;
      par[0][3] = u_jet_;
      par[0][4] = T_jet_;
      par[0][5] = u_far_;
      par[0][6] = T_far_;
      par[0][11] = p_jet_;
      par[0][12] = super_sonic_ ? 1.0 : 0.0;
    } else {
      bc[0] = DRN_BC_SLIP;
    }
    for (int f = 2; f < 6; ++f) {
      bc[f] = DRN_BC_FARFIELD;
      par[f][0] = 1e5;  // the far-field functor's default state, at rest
      par[f][1] = 300.0;
    }
    return true;
  }

  double pFar() const { return p_far_; }

 private:
  double u_jet_ = 0, u_far_ = 0, p_jet_ = 0, p_far_ = 0, T_jet_ = 0, T_far_ = 0, radius_ = 0;
  int jet_patch_ = -1;
  bool super_sonic_ = false;
};

}  // namespace drnum
