"""The jet demo's flux operator through the C++ drop-in (drnum_amd/host/jetflux.h):
JetFlux chooses the face functors of every patch in its wall methods
(applications/drnumJetDemo/jetflux.h:140-198) -- the top-hat jet on the jet
patch's x-minus face, slip walls on the other x faces, the far-field functor
with its default state (1e5 Pa, 300 K, at rest) on the y and z faces.  The
drop-in takes those choices from the operator (its faces() member) instead of
the grid's face codes.

  * drnum_run with jetflux= and a case file whose face codes are all
    transmissive gives the oracle's result for the faces JetFlux chooses
    (FP64 1e-11, FP32 2e-5), in the reference's stage loop and in one library
    call per step;
  * it is bit-identical to drnum_run without jetflux= on a case file that
    carries JetFlux's face codes itself.
"""
import numpy as np
import pytest

from drnum_amd import configs, gas

from test_gpu_dropin import oracle_after, run_app
from test_gpu_parity import TOL, rel_linf

pytestmark = pytest.mark.gpu

JET = dict(patch=0, u_jet=100.0, u_far=10.0, p_jet=1e5, p_far=1e5, T_jet=300.0, T_far=300.0, radius=0.2)


def jetflux_arg():
    j = JET
    return "jetflux=%d,%r,%r,%r,%r,%r,%r,%r,0" % (j["patch"], j["u_jet"], j["u_far"], j["p_jet"], j["p_far"],
                                                   j["T_jet"], j["T_far"], j["radius"])


def jet_demo_workload(faces_from_jetflux):
    """The 8-patch jet grid at rest at 1e5 Pa and 300 K.  With
    faces_from_jetflux every patch face carries the functor JetFlux chooses for
    it; otherwise every face is transmissive (a case file the operator overrides)."""
    w = configs.jet(core=8, blocks=2)
    for p, P in enumerate(w.patches):
        par = np.zeros((6, configs.NPAR))
        if faces_from_jetflux:
            bc = [configs.BC_SLIP, configs.BC_SLIP] + [configs.BC_FARFIELD] * 4
            for f in range(2, 6):
                par[f, :5] = [1e5, 300.0, 0.0, 0.0, 0.0]
            if p == JET["patch"]:
                bc[0] = configs.BC_JET
                par[0, 0] = 0.5 * P.nj * P.dx(1)
                par[0, 1] = 0.5 * P.nk * P.dx(2)
                par[0, 2] = JET["radius"]
                par[0, 3] = JET["u_jet"]
                par[0, 4] = JET["T_jet"]
                par[0, 5] = JET["u_far"]
                par[0, 6] = JET["T_far"]
                par[0, 11] = JET["p_jet"]
        else:
            bc = [configs.BC_TRANSMISSIVE] * 6
        P.bc = bc
        P.bc_par = par

    def init(x, y, z):
        one = np.ones_like(x)
        return np.stack([1e5 * one, 300.0 * one, 0 * x, 0 * x, 0 * x])
    w.init = init
    a = np.sqrt(1.4 * gas.R * 300.0)
    w.dt = 0.3 * (1.0 / 16) / (a + JET["u_jet"])
    return w


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("mode", ["rk", "hip"])
def test_jetflux_chooses_the_faces(gpu, tmp_path, dtype, mode):
    steps = 5
    got = run_app(tmp_path, jet_demo_workload(False), dtype, steps, mode, extra=[jetflux_arg()], tag="_op")
    ref_w = jet_demo_workload(True)
    og = oracle_after(ref_w, dtype, steps)
    err = rel_linf(got, og)
    print("JetFlux operator", mode, np.dtype(dtype).name, "vs oracle", err, flush=True)
    assert err <= TOL[dtype]
    # the jet has started: mass has entered through the jet patch's x-minus face
    assert float(got[JET["patch"]][1].max()) > 0.0
    same = run_app(tmp_path, ref_w, dtype, steps, mode, tag="_codes")
    for a, b in zip(got, same):
        assert np.array_equal(a, b)
