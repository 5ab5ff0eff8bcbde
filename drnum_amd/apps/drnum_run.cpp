// drnum_run.cpp -- a DrNUM-style application on the HIP drop-in.
//
// Written the way the reference's demo mains drive the path
// (applications/drnumJetDemo/main.h:90-240): read a grid file, compute the
// donor dependencies, set the initial field, build the iterator for the flux
// operator AusmPlus<5, Upwind2<5, VanAlbada>, PerfectGas>, add the RK
// coefficients, run the time steps, copy the result back to the host.
//
// usage: drnum_run <grid file> <case file> <out file> <steps> <dt> <mode> [f64] [levelset] [swall] [visc]
//                  [restart=<file.dnd>] [dnd=<base name>] [codes=<code string>] [vtk=<base name>]
//                  [cfl=<target>]  (dt rescaled to the target CFL number after each step)
//                  [jetflux=<parameters>] [flags=<file>]
//   case file (binary, little endian), per patch in grid order:
//     int32 bc[6], float64 par[6][16], then 5*N initial values in `real`
//     (variable-major, the layout of field 0 in Patch::m_Data);
//     with `levelset`, followed by int32 n_spheres, int32 n_inner,
//     float64 min_inner_rel and n_spheres x (xc, yc, zc, r) in float64
//   visc: the operator is EulerViscFlux<AusmPlus<...>, CompressibleViscFlux<5, PerfectGas>>
//   mode: rk     RungeKutta over the virtual iterator interface (reference loop)
//         hip    HIP_RungeKutta (one library call per step)
//         graph  HIP_RungeKutta::graph (captured steps)
//   out file: field 0 of every patch, 5*N values in `real` each
//   restart=F: field 0 is read from the DrNUM restart file F after the case
//              file (PatchGrid::readData), the time from its header
//   dnd=B:     field 0 is also written to B_<steps, 6 digits>.dnd at the end
//              (PatchGrid::writeData), with time = restart time + steps*dt
//   codes=S:   the iterator gets the code string S (commas may separate the
//              words) and an IteratorFeeder hands it the patch groups of equal
//              code string; a group it does not match stops the run (exit 5).
//              Without codes=, every patch is added to the iterator.
//   vtk=B:     at the end, B_<steps, 6 digits>.vtm and B_<steps>/ with one .vts
//              per patch (PatchGrid::writeToVtk with CompressibleVariables)
//   jetflux=P,u_jet,u_far,p_jet,p_far,T_jet,T_far,radius,supersonic:
//              the operator is the jet demo's JetFlux (drnum_amd/host/jetflux.h,
//              set up as JetFlux::setup, jetflux.h:86-97) around the Euler flux
//              (or the viscous one with `visc`); it chooses every patch's face
//              functors (the jet on patch P's x-minus face), in place of the
//              case file's codes
//   flags=F:   per-cell flags, one byte per cell of every patch in grid order
//              (bit 0 inactive, bit 1 inside): inactive cells are deactivated
//              (Patch::deactivate) and the inside cells given to the patch as
//              split faces (Patch::setSplitFaces: every face between an inside
//              and an outside cell, from both sides), from which the patch
//              fills its inside
// Prints one line per phase and exits non-zero on any error.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../host/iterators.h"
#include "../host/jetflux.h"

using namespace drnum;

typedef AusmPlus<5, Upwind2<5, VanAlbada>, PerfectGas> euler_t;
// the same Euler flux plus the viscous flux on every interior face (EaFlux,
// drnumCylinder/main.h:127-163)
typedef EulerViscFlux<euler_t, CompressibleViscFlux<5, PerfectGas>> visc_t;

struct Options {
  std::string grid_file, case_file, out_file, mode, restart, dnd, codes, vtk, jetflux, flags;
  int steps = 0;
  double dt = 0;
  bool levelset = false;
  bool swall = false;  // the level-set fill with LSOBCCompressibleSWallOp instead of the slip op
  double cfl = 0;      // > 0: dt rescaled to this CFL number after every step (rk and hip modes)
};

template <typename real, class OP>
static int run(const Options& o, const OP& op = OP()) {
  const std::string &grid_file = o.grid_file, &case_file = o.case_file, &out_file = o.out_file, &mode = o.mode;
  const std::string &restart = o.restart, &dnd = o.dnd, &codes = o.codes, &vtk = o.vtk;
  const int steps = o.steps;
  double dt = o.dt;
  double t_run = 0;  // simulated time of this run
  const bool levelset = o.levelset;
  PatchGrid<double> grid;
  std::string err;
  if (grid.readGrid(grid_file, 1.0, &err)) {
    fprintf(stderr, "readGrid: %s\n", err.c_str());
    return 2;
  }
  std::ifstream in(case_file, std::ios::binary);
  if (!in) {
    fprintf(stderr, "cannot open %s\n", case_file.c_str());
    return 2;
  }
  std::vector<HostPatch<real>> patches;
  patches.reserve(grid.numPatches());
  for (size_t p = 0; p < grid.numPatches(); ++p) patches.emplace_back(&grid, p);
  for (size_t p = 0; p < grid.numPatches(); ++p) {
    CartesianPatch<double>& P = grid.patch(p);
    int32_t bc[6];
    in.read(reinterpret_cast<char*>(bc), sizeof(bc));
    in.read(reinterpret_cast<char*>(P.face_par), sizeof(P.face_par));
    for (int f = 0; f < 6; ++f) P.face[f] = bc[f];
    in.read(reinterpret_cast<char*>(patches[p].getField(0)), (std::streamsize)(5 * P.cells() * sizeof(real)));
  }
  int n_spheres = 0, n_inner = 0;
  double min_rel = 0;
  std::vector<double> spheres;
  if (levelset) {
    int32_t ns = 0, ni = 0;
    in.read(reinterpret_cast<char*>(&ns), 4);
    in.read(reinterpret_cast<char*>(&ni), 4);
    in.read(reinterpret_cast<char*>(&min_rel), 8);
    n_spheres = ns;
    n_inner = ni;
    spheres.resize(4 * (size_t)n_spheres);
    in.read(reinterpret_cast<char*>(spheres.data()), (std::streamsize)(spheres.size() * sizeof(double)));
  }
  if (!in) {
    fprintf(stderr, "case file %s is short\n", case_file.c_str());
    return 2;
  }
  const double period[3] = {0, 0, 0};
  if (grid.computeDependencies(period, 8)) {
    fprintf(stderr, "computeDependencies hit a BUG condition\n");
    return 2;
  }
  if (levelset && grid.levelSetSpheres(spheres, n_inner, 0, min_rel)) {
    fprintf(stderr, "level-set mirror search hit a BUG condition\n");
    return 2;
  }
  printf("grid: %zu patches read, dependencies computed\n", grid.numPatches());
  std::vector<HostPatch<real>*> pp;
  for (auto& P : patches) pp.push_back(&P);
  double t0 = 0;
  if (!restart.empty()) {
    t0 = (double)readData<real>(pp, 0, restart);
    printf("restart: t = %g from %s\n", t0, restart.c_str());
  }

  if (!o.flags.empty()) {
    std::ifstream fin(o.flags, std::ios::binary);
    size_t n_inactive = 0, n_faces = 0, n_inside = 0;
    for (size_t p = 0; p < grid.numPatches(); ++p) {
      const CartesianPatch<double>& P = grid.patch(p);
      std::vector<uint8_t> f(P.cells());
      fin.read(reinterpret_cast<char*>(f.data()), (std::streamsize)f.size());
      std::vector<splitface_t> faces;
      for (size_t l = 0; l < P.cells(); ++l) {
        if (f[l] & DRN_CELL_INACTIVE) {
          patches[p].deactivate(l);
          ++n_inactive;
        }
        size_t i, j, k;
        P.ijk(l, i, j, k);
        const size_t up[3][3] = {{i + 1, j, k}, {i, j + 1, k}, {i, j, k + 1}};
        for (const auto& q : up) {
          if (!P.inRange(q[0], q[1], q[2])) continue;
          const size_t m = P.index(q[0], q[1], q[2]);
          if (((f[l] ^ f[m]) & DRN_CELL_INSIDE) == 0) continue;
          splitface_t a, b;
          a.idx = l;
          a.idx_neigh = m;
          a.inside = (f[l] & DRN_CELL_INSIDE) != 0;
          b.idx = m;
          b.idx_neigh = l;
          b.inside = (f[m] & DRN_CELL_INSIDE) != 0;
          faces.push_back(a);
          faces.push_back(b);
        }
      }
      patches[p].setSplitFaces(faces);
      n_faces += faces.size() / 2;
      for (size_t l = 0; l < P.cells(); ++l) n_inside += patches[p].isInsideCell(l) ? 1 : 0;
    }
    if (!fin) {
      fprintf(stderr, "flags file %s is short\n", o.flags.c_str());
      return 2;
    }
    printf("flags: %zu inactive cells, %zu split faces, %zu inside cells\n", n_inactive, n_faces, n_inside);
  }

  HIP_CartesianIterator<5, OP, real> iterator(op, 0);
  if (codes.empty()) {
    for (auto& P : patches) iterator.addPatch(&P);
  } else {
    iterator.setCodeString(CodeString(codes));
    IteratorFeeder<real> feeder;
    feeder.addIterator(&iterator);
    try {
      feeder.feed(grid, pp);
    } catch (const FeedError& e) {
      fprintf(stderr, "feed: %s\n", e.what());
      return 5;
    }
    printf("feed: %zu patches for \"%s\"\n", iterator.numPatches(), iterator.getCodeString().c_str());
  }
  iterator.updateDevice();
  GenericOperation* ls = nullptr;
  if (levelset && o.swall)
    ls = new HIP_LevelSetBC<5, OP, real, LSOBCCompressibleSWallOp>(&iterator);
  else if (levelset)
    ls = new HIP_LevelSetBC<5, OP, real>(&iterator);

  RungeKutta<real> rk;
  HIP_RungeKutta<5, OP, real> hrk(&iterator);
  for (real a : {real(0.25), real(0.5), real(1.0)}) {
    rk.addAlpha(a);
    hrk.addAlpha(a);
  }
  rk.addIterator(&iterator);
  if (ls) {
    rk.addPostOperation(ls);
    hrk.addPostOperation(ls);
  }

  // with cfl=<target>, dt *= target / CFL_max after every step, the rescale of
  // the jet demo (applications/drnumJetDemo/main.h:233), which applies it at
  // every write interval
  auto rescale = [&]() {
    if (o.cfl <= 0) return;
    const double c = iterator.cflMax(0, dt);
    if (c > 0) dt *= o.cfl / c;
    printf("dt %.12e CFL %.12e\n", dt, c);
  };
  if (o.cfl > 0 && mode == "graph") {
    fprintf(stderr, "cfl= needs mode rk or hip (a graph replays one dt)\n");
    return 2;
  }
  if (mode == "rk") {
    for (int s = 0; s < steps; ++s) {
      rk(real(dt));
      t_run += dt;
      rescale();
    }
  } else if (mode == "hip") {
    for (int s = 0; s < steps; ++s) {
      hrk(real(dt));
      t_run += dt;
      rescale();
    }
  } else if (mode == "graph") {
    hrk.graph(real(dt), steps);
    t_run = steps * dt;
  } else {
    fprintf(stderr, "unknown mode %s\n", mode.c_str());
    return 2;
  }
  iterator.sync();
  printf("ran %d steps (%s)\n", steps, mode.c_str());
  {
    // the density change of the last step, as the jet demo logs it
    // (applications/drnumJetDemo/main.h:216-225): max over patches, L2 as the
    // root mean square of the patches' L2 norms
    real max_all = 0, ql2 = 0;
    for (size_t p = 0; p < iterator.numPatches(); ++p) {
      real mx = 0, l2 = 0;
      iterator.computeVariableDifference(p, 0, 0, 1, 0, mx, l2);
      max_all = std::max(max_all, mx);
      ql2 += l2 * l2;
    }
    printf("drho: max %.6e L2 %.6e\n", (double)max_all, std::sqrt((double)ql2 / std::max<size_t>(1, iterator.numPatches())));
  }
  iterator.updateHost();
  std::ofstream out(out_file, std::ios::binary);
  for (auto& P : patches)
    out.write(reinterpret_cast<const char*>(P.getField(0)), (std::streamsize)(5 * P.variableSize() * sizeof(real)));
  delete ls;
  printf("wrote %s\n", out_file.c_str());
  if (!dnd.empty()) {
    writeData<real>(pp, 0, dnd, real(t0 + t_run), steps);
    printf("wrote restart file %s_%06d.dnd\n", dnd.c_str(), steps);
  }
  if (!vtk.empty()) {
    CompressibleVariables<PerfectGas> vars;
    iterator.writeToVtk(0, vtk, vars, steps);
    printf("wrote %s_%06d.vtm\n", vtk.c_str(), steps);
  }
  return out ? 0 : 3;
}

int main(int argc, char** argv) {
  if (argc < 7) {
    fprintf(stderr,
            "usage: %s <grid> <case> <out> <steps> <dt> <rk|hip|graph> [f64] [levelset] [swall] [visc] [restart=F] [dnd=B] "
            "[codes=S] [vtk=B] [cfl=C] [jetflux=...] [flags=F]\n",
            argv[0]);
    return 2;
  }
  Options o;
  o.grid_file = argv[1];
  o.case_file = argv[2];
  o.out_file = argv[3];
  o.steps = atoi(argv[4]);
  o.dt = atof(argv[5]);
  o.mode = argv[6];
  bool f64 = false, visc = false;
  for (int a = 7; a < argc; ++a) {
    if (!strcmp(argv[a], "f64")) f64 = true;
    if (!strcmp(argv[a], "levelset")) o.levelset = true;
    if (!strcmp(argv[a], "swall")) o.swall = true;
    if (!strncmp(argv[a], "cfl=", 4)) o.cfl = atof(argv[a] + 4);
    if (!strcmp(argv[a], "visc")) visc = true;
    if (!strncmp(argv[a], "restart=", 8)) o.restart = argv[a] + 8;
    if (!strncmp(argv[a], "dnd=", 4)) o.dnd = argv[a] + 4;
    if (!strncmp(argv[a], "codes=", 6)) o.codes = argv[a] + 6;
    if (!strncmp(argv[a], "vtk=", 4)) o.vtk = argv[a] + 4;
    if (!strncmp(argv[a], "jetflux=", 8)) o.jetflux = argv[a] + 8;
    if (!strncmp(argv[a], "flags=", 6)) o.flags = argv[a] + 6;
  }
  try {
    if (!o.jetflux.empty()) {
      double v[9];
      if (sscanf(o.jetflux.c_str(), "%lf,%lf,%lf,%lf,%lf,%lf,%lf,%lf,%lf", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5],
                 &v[6], &v[7], &v[8]) != 9) {
        fprintf(stderr, "jetflux= needs 9 comma-separated values\n");
        return 2;
      }
      if (visc) {
        JetFlux<visc_t> op;
        op.setup(v[1], v[2], v[3], v[4], v[5], v[6], v[7], (int)v[0], v[8] != 0);
        return f64 ? run<double>(o, op) : run<float>(o, op);
      }
      JetFlux<euler_t> op;
      op.setup(v[1], v[2], v[3], v[4], v[5], v[6], v[7], (int)v[0], v[8] != 0);
      return f64 ? run<double>(o, op) : run<float>(o, op);
    }
    if (f64 && visc) return run<double, visc_t>(o);
    if (f64) return run<double, euler_t>(o);
    if (visc) return run<float, visc_t>(o);
    return run<float, euler_t>(o);
  } catch (const DeviceError& e) {
    fprintf(stderr, "device error: %s\n", e.what());
    return 4;
  }
}
