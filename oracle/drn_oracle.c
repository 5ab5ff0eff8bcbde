/*
 * drn_oracle.c -- TEST INFRASTRUCTURE ONLY.  A plain-C restatement of the
 * DrNUM per-patch explicit update (flux/residual, RK stage update, donor
 * exchange, level-set fill) and of the host-side table builders that feed it
 * (donor tables, level-set layers).  It is the checker for the HIP path and
 * the timed CPU baseline; the product never links or calls it.
 *
 * Compiled twice (see Makefile): O_REAL=float -> o32_*, O_REAL=double -> o64_*.
 * Arithmetic follows the reference expression by expression, including the
 * places where a double literal promotes a float expression to double
 * (e.g. perfectgas.h:63, compressiblefarfieldflux.h:101), so that the float
 * build reproduces the reference's default (float) build and the double build
 * its FP64 branch.  Build with -ffp-contract=off and without -ffast-math.
 *
 * Reference files restated (paths under src/drnumlib/ unless stated):
 *   perfectgas.h, fluxes/compressibleflux.h, fluxes/ausmplus.h,
 *   fluxes/compressibleslipflux.h, fluxes/compressiblefarfieldflux.h,
 *   reconstruction/{upwind1,upwind2,vanalbada,minmod,secondorder}.h, drnum.h,
 *   iterators/cartesianiterator.h, rungekutta.cpp, timeintegration.cpp,
 *   cartesianpatch.cpp, patch.cpp, patchgrid.cpp, weightedset.h,
 *   intercoeffws.h, intercoeffpad.h, math/coordtransform*.{h,cpp},
 *   levelsetobject.cpp, spherelevelset.cpp, combilevelsetor.cpp,
 *   cpu_levelsetobjectbc.h, lsobccompressibleeulerop.h, levelsetobjectbc.cpp,
 *   applications/drnumJetDemo/jetflux.h, fluxes/compressibleviscflux.h,
 *   cartesianpatch_common.h (getXGrad/getYGrad/getZGrad),
 *   applications/drnumCylinder/main.h (EaFlux: Euler flux, then viscous flux).
 *
 * Patches may be turned (base vectors, patch.cpp:63-85).  Per-cell flags
 * (grid_set_cell_flags) mark inactive cells, which the update skips
 * (cartesianiterator.h:315), and inside cells, between which and outside
 * cells a face is a split face where MUSCL falls back to first order
 * (upwind2.h:41); without flags every cell is active and none is inside, as
 * patch.cpp:703-715 allocates them.
 */
#include "drn_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define F(n) CAT(O_PFX, n)

typedef O_REAL REAL;
typedef struct O_GRID GRID;

#if O_IS_FLOAT
#define RSQRT sqrtf
#define RFABS fabsf
#else
#define RSQRT sqrt
#define RFABS fabs
#endif

/* perfectgas.h:31-34 -- the literals are double and converted to real */
#define GAS_R ((REAL)287.2)
#define GAS_GAMMA ((REAL)1.4)
#define GAS_CP ((REAL)1004.5)
#define GAS_CV ((REAL)717.5)
/* drnum.h:65,68,70 */
#define FR12 ((REAL)0.5)
#define FR23 ((REAL)0.6666666)
#define GLOBAL_EPS ((REAL)1e-10)

/* ------------------------------------------------------------------------ */
/* patch and grid                                                             */
/* ------------------------------------------------------------------------ */

typedef struct {
  size_t n;
  size_t idx[8];
  REAL w[8];
} wset_t;

typedef struct {
  int patch;         /* donor patch                                      */
  size_t nrec;       /* receiver cells served by this donor              */
  size_t stride;     /* padded contributions per receiver                */
  size_t rec_start;  /* start in the concatenated receiver list          */
  size_t wi_start;   /* start in the concatenated index/weight lists     */
  REAL a[3][3];      /* donor -> receiver turn of vector variables (donor_t.h:36-44) */
} donor_t_;

/* CoordTransform (math/coordtransform.h:26-94): x -> A x + b */
typedef struct {
  REAL a[3][3];
  REAL b[3];
} ctrans_t;

typedef struct {
  size_t cell;
  REAL g, gx, gy, gz;
  int ex_ok;
  size_t donor[8];
  REAL weight[8];
} lscell_t;

typedef struct {
  size_t n, cap;
  lscell_t* v;
} lsvec_t;

typedef struct {
  int ni, nj, nk;
  size_t N, njk;
  int seek[6];
  REAL o[3];
  REAL base_i[3], base_j[3]; /* as read (patch.cpp:68-76) */
  ctrans_t i2t, t2i;  /* m_TransformInertial2This: inertial -> this and its stored inverse */
  REAL L[3];
  REAL dx, dy, dz;
  REAL xccmin, yccmin, zccmin;
  size_t dzf[3], dza[3]; /* donor zone [first, afterlast) per axis */
  int bc[6];
  double bc_par[6][O_BC_NPAR];
  REAL far_var[6][5];
  REAL* data; /* 3 fields x 5 vars x N */
  unsigned char* flags; /* per cell: O_CELL_INACTIVE, O_CELL_INSIDE; NULL = all active, none inside */

  /* dependencies (patch.cpp:128-218 direct transfer lists) */
  size_t* recv;
  size_t n_recv;
  size_t n_no_donor;
  int n_donors;
  donor_t_* donors;
  size_t n_rec_concat;
  size_t* rec_concat;
  size_t n_wi;
  size_t* donor_idx;
  REAL* donor_w;
  int64_t* recv64;
  int64_t* rec_concat64;
  int64_t* donor_idx64;

  /* level set */
  REAL* G;
  int n_inner, n_outer;
  lsvec_t* inner;
  lsvec_t* outer;
} opatch;

struct O_GRID {
  int np;
  opatch* p;
  int recon, lim;
  int visc; /* add CompressibleViscFlux to every interior face flux */
  int nthreads;
  int have_ls;
  int ls_op;  /* 0: LSOBCCompressibleEulerOp, 1: LSOBCCompressibleSWallOp */
  int error;
};

static size_t IDX(const opatch* P, size_t i, size_t j, size_t k) { return i * P->njk + j * (size_t)P->nk + k; }

/* cartesianpatch_common.h:306-313 (unsigned check: negative wraps to huge) */
static int in_range(const opatch* P, size_t i, size_t j, size_t k) {
  return !(i >= (size_t)P->ni || j >= (size_t)P->nj || k >= (size_t)P->nk);
}

static REAL* fieldp(const opatch* P, int field) { return P->data + (size_t)field * 5 * P->N; }

/* ------------------------------------------------------------------------ */
/* gas model                                                                  */
/* ------------------------------------------------------------------------ */

/* perfectgas.h:56-64 (var[4] is evaluated in double: 0.5 is a double literal) */
void F(prim_to_cons)(REAL p, REAL T, REAL u, REAL v, REAL w, REAL* var) {
  var[0] = p / (GAS_R * T);
  var[1] = var[0] * u;
  var[2] = var[0] * v;
  var[3] = var[0] * w;
  var[4] = var[0] * (GAS_CV * T + 0.5 * (u * u + v * v + w * w));
}

/* perfectgas.h:83-92 */
static void cons_to_prim(const REAL* var, REAL* p, REAL* T, REAL* u, REAL* v, REAL* w) {
  REAL ir = 1.0 / var[0];
  *u = var[1] * ir;
  *v = var[2] * ir;
  *w = var[3] * ir;
  *T = (var[4] * ir - 0.5 * ((*u) * (*u) + (*v) * (*v) + (*w) * (*w))) / GAS_CV;
  *p = var[0] * GAS_R * (*T);
}

/* ------------------------------------------------------------------------ */
/* reconstruction                                                             */
/* ------------------------------------------------------------------------ */

/* drnum.h:256-263 with std::min / std::max semantics */
static REAL nonzero(REAL x, REAL eps) {
  if (x < 0) {
    REAL a = -eps;
    return (x < a) ? x : a;
  }
  return (eps < x) ? x : eps;
}

/* vanalbada.h:28-33, minmod.h:28-31, secondorder.h:28-31 */
REAL F(limiter)(int lim, REAL d1, REAL d2) {
  if (lim == O_LIM_VANALBADA) {
    REAL r = d2 / nonzero(d1, GLOBAL_EPS);
    return (r * r + r) / (r * r + 1);
  } else if (lim == O_LIM_MINMOD) {
    REAL t = (REAL)1 * d2 / nonzero(d1, GLOBAL_EPS);
    REAL m = (t < (REAL)1) ? t : (REAL)1;
    return ((REAL)0 < m) ? m : (REAL)0;
  }
  return 1;
}

/* Patch::isSplitFace (patch_common.h:500-503): the face between two cells one
 * of which is inside (m_IsInsideCell) and the other not */
static int split_face(const opatch* P, size_t c0, size_t c1) {
  if (!P->flags) return 0;
  return ((P->flags[c0] & O_CELL_INSIDE) != 0) != ((P->flags[c1] & O_CELL_INSIDE) != 0);
}

/* upwind2.h:29-52 and upwind1.h:29-39: second order only where i0 and i2 lie in
 * the patch and the face i0-i1 is not a split face */
static void project(const opatch* P, const REAL* f, int recon, int lim, size_t i1, size_t j1, size_t k1,
                    size_t i2, size_t j2, size_t k2, REAL* var) {
  const size_t N = P->N;
  const size_t c1 = IDX(P, i1, j1, k1);
  for (int v = 0; v < 5; ++v) var[v] = f[v * N + c1];
  if (recon == O_RECON_UPWIND1) return;
  size_t i0 = 2 * i1 - i2, j0 = 2 * j1 - j2, k0 = 2 * k1 - k2;
  if (in_range(P, i0, j0, k0) && in_range(P, i2, j2, k2) && !split_face(P, IDX(P, i0, j0, k0), c1)) {
    const size_t c0 = IDX(P, i0, j0, k0), c2 = IDX(P, i2, j2, k2);
    for (int v = 0; v < 5; ++v) {
      REAL d01 = f[v * N + c1] - f[v * N + c0];
      REAL d12 = f[v * N + c2] - f[v * N + c1];
      var[v] += (REAL)0.5 * F(limiter)(lim, d01, d12) * d01;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* AUSM+                                                                      */
/* ------------------------------------------------------------------------ */

typedef struct {
  REAL r, u, v, w, T, p, a, H;
} prim_t;

/* compressibleflux.h:236-268 (COMPRESSIBLE_LEFT_VARS / RIGHT_VARS) */
static prim_t lr_vars(const REAL* var) {
  prim_t s;
  REAL r = var[0];
  REAL ir = 1.0 / r;
  s.u = var[1] * ir;
  s.v = var[2] * ir;
  s.w = var[3] * ir;
  REAL rE = var[4];
  s.T = (rE * ir - (REAL)0.5 * (s.u * s.u + s.v * s.v + s.w * s.w)) / GAS_CV;
  s.p = r * GAS_R * s.T;
  s.a = RSQRT(GAS_GAMMA * GAS_R * s.T);
  s.H = (rE + s.p) / r;
  s.r = r;
  return s;
}

/* compressibleflux.h:46-75 */
static REAL M1f(REAL M, REAL s) { return (REAL)(0.5 * (M + s * RFABS(M))); }
static REAL M2f(REAL M, REAL s) {
  REAL t = M + s;
  REAL sq = t * t;
  return (REAL)(0.25 * s * sq);
}
static REAL M4f(REAL M, REAL s) {
  if (RFABS(M) >= 1) return M1f(M, s);
  return M2f(M, s) * (1 - 2 * s * M2f(M, -s));
}
static REAL P5f(REAL M, REAL s) {
  if (RFABS(M) >= 1) return M1f(M, s) / M;
  return M2f(M, s) * ((2 * s - M) - 3 * s * M * M2f(M, -s));
}

static REAL comp3(int d, REAL u, REAL v, REAL w) { return d == 0 ? u : (d == 1 ? v : w); }

/* ausmplus.h:39-106 (xField/yField/zField differ only in the normal component) */
void F(ausm_flux)(const REAL* var_l, const REAL* var_r, int dir, REAL A, REAL* flux) {
  prim_t L = lr_vars(var_l), R = lr_vars(var_r);
  REAL unl = comp3(dir, L.u, L.v, L.w), unr = comp3(dir, R.u, R.v, R.w);
  REAL a = FR12 * (L.a + R.a);
  REAL M = M4f(unl / a, 1) + M4f(unr / a, -1);
  REAL Mp = FR12 * (M + RFABS(M));
  REAL Mm = FR12 * (M - RFABS(M));
  REAL p = P5f(unl / a, 1) * L.p + P5f(unr / a, -1) * R.p;
  flux[0] += a * A * (L.r * Mp + R.r * Mm);
  REAL f0 = flux[0];
  REAL vl[3] = {L.u, L.v, L.w}, vr[3] = {R.u, R.v, R.w};
  for (int c = 0; c < 3; ++c) {
    if (c == dir)
      flux[1 + c] += FR12 * f0 * (vl[c] + vr[c]) + A * p - FR12 * RFABS(f0) * (vr[c] - vl[c]);
    else
      flux[1 + c] += FR12 * f0 * (vl[c] + vr[c]) - FR12 * RFABS(f0) * (vr[c] - vl[c]);
  }
  flux[4] += FR12 * f0 * (L.H + R.H) - FR12 * RFABS(f0) * (R.H - L.H);
}

/* ------------------------------------------------------------------------ */
/* wall functors                                                              */
/* ------------------------------------------------------------------------ */

/* compressibleslipflux.h:34-98 with COMPR_VARS (compressibleflux.h:26-38) */
static void wall_slip(const REAL* var, int dir, REAL A, REAL* flux) {
  REAL r = var[0];
  REAL ir = (REAL)1.0 / r;
  REAL u = var[1] * ir, v = var[2] * ir, w = var[3] * ir;
  REAL rE = var[4];
  REAL T = (rE * ir - (REAL)0.5 * (u * u + v * v + w * w)) / GAS_CV;
  REAL p = r * GAS_R * T;
  flux[1 + dir] += A * p;
}

/* Not in the reference: zero-gradient outflow, the physical Euler flux of the
 * boundary cell state (mass A*rho*u_n, momentum + A*p on the normal, energy A*rho*u_n*H). */
static void wall_transmissive(const REAL* var, int dir, REAL A, REAL* flux) {
  REAL r = var[0];
  REAL ir = (REAL)1.0 / r;
  REAL u = var[1] * ir, v = var[2] * ir, w = var[3] * ir;
  REAL rE = var[4];
  REAL T = (rE * ir - (REAL)0.5 * (u * u + v * v + w * w)) / GAS_CV;
  REAL p = r * GAS_R * T;
  REAL m = A * var[1 + dir];
  REAL vel[3] = {u, v, w};
  flux[0] += m;
  for (int c = 0; c < 3; ++c) {
    if (c == dir)
      flux[1 + c] += m * vel[c] + A * p;
    else
      flux[1 + c] += m * vel[c];
  }
  flux[4] += m * ((rE + p) * ir);
}

/* shared tail of compressiblefarfieldflux.h:101-105 / jetflux.h:186-190 */
static void wall_state_flux(REAL p, REAL T, const REAL* vel, int dir, REAL A, REAL* flux) {
  REAL un = vel[dir];
  flux[0] = A * p / (GAS_R * T) * un;
  for (int c = 0; c < 3; ++c) {
    if (c == dir)
      flux[1 + c] = flux[0] * vel[c] + A * p;
    else
      flux[1 + c] = flux[0] * vel[c];
  }
  flux[4] = flux[0] * (GAS_CP * T + 0.5 * (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]));
}

/* compressiblefarfieldflux.h:26-65,93-181 (Upwind1 projection: var1 is the boundary cell) */
static void wall_farfield(const REAL* far_var, const REAL* var1, int dir, int plus, REAL A, REAL* flux) {
  REAL p0, T0, u0, v0, w0, p1, T1, u1, v1, w1;
  cons_to_prim(far_var, &p0, &T0, &u0, &v0, &w0);
  cons_to_prim(var1, &p1, &T1, &u1, &v1, &w1);
  REAL a0 = RSQRT(GAS_GAMMA * GAS_R * T0);
  REAL a1 = RSQRT(GAS_GAMMA * GAS_R * T1);
  REAL M0 = comp3(dir, u0, v0, w0) / a0;
  REAL M1 = comp3(dir, u1, v1, w1) / a1;
  if (plus) {
    M0 = -M0;
    M1 = -M1;
  }
  REAL M = (REAL)0.5 * (M0 + M1);
  REAL p, T, vel[3];
  if (M >= 0) {
    if (M >= 1) {
      p = p0;
    } else {
      p = p1;
    }
    T = T0;
    vel[0] = u0; vel[1] = v0; vel[2] = w0;
  } else {
    /* the `M <= 1` test of FARFIELD_SPLIT is always true for M < 0 */
    p = p1;
    T = T1;
    vel[0] = u1; vel[1] = v1; vel[2] = w1;
  }
  wall_state_flux(p, T, vel, dir, A, flux);
}

static uint64_t smix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

/* counter-based uniform noise in [-1,1) for inflow face cell (ga, gb), component c */
static REAL jet_noise(uint64_t seed, int64_t ga, int64_t gb, int c) {
  uint64_t key = ((uint64_t)(uint32_t)ga << 32) | (uint64_t)(uint32_t)gb;
  uint64_t base = smix64(key ^ smix64(seed));
  uint64_t h = smix64(base + (uint64_t)c);
  int64_t q = (int64_t)(h >> 40) - 8388608;
  return (REAL)q * (REAL)(1.0 / 8388608.0);
}

/* jetflux.h:158-192 (JetFlux::xWallM), generalised to any face: inside the jet
 * radius the normal velocity is the jet speed (with seeded velocity noise) and
 * p is p_jet when the jet is supersonic (par[12] != 0), else the boundary
 * cell's; outside the radius the normal velocity is the far speed and p the
 * cell's. */
static void wall_jet(const double* par, const REAL* var, int dir, int plus, REAL ta, REAL tb, int64_t la,
                     int64_t lb, REAL A, REAL* flux) {
  REAL p, T, u, v, w;
  cons_to_prim(var, &p, &T, &u, &v, &w);
  REAL ca = (REAL)par[0], cb = (REAL)par[1], rad = (REAL)par[2];
  REAL rr = (ta - ca) * (ta - ca) + (tb - cb) * (tb - cb);
  REAL sgn = plus ? (REAL)-1 : (REAL)1;
  REAL vel[3] = {0, 0, 0};
  int axa = dir == 0 ? 1 : 0;
  int axb = dir == 2 ? 1 : 2;
  if (rr < rad * rad) {
    REAL uj = (REAL)par[3], amp = (REAL)par[7];
    uint64_t seed = (uint64_t)par[8];
    int64_t ga = la + (int64_t)par[9], gb = lb + (int64_t)par[10];
    REAL na = jet_noise(seed, ga, gb, 0), nb = jet_noise(seed, ga, gb, 1), nn = jet_noise(seed, ga, gb, 2);
    vel[dir] = sgn * (uj * (1 + amp * nn));
    vel[axa] = uj * (amp * na);
    vel[axb] = uj * (amp * nb);
    if (par[12] != 0.0) p = (REAL)par[11]; /* jetflux.h:176-178 (m_SuperSonic) */
    T = (REAL)par[4];
  } else {
    vel[dir] = sgn * (REAL)par[5];
    T = (REAL)par[6];
  }
  wall_state_flux(p, T, vel, dir, A, flux);
}

/* ------------------------------------------------------------------------ */
/* iterator: residual + RK stage update                                       */
/* ------------------------------------------------------------------------ */

static void get_var(const opatch* P, const REAL* f, size_t i, size_t j, size_t k, REAL* var) {
  size_t c = IDX(P, i, j, k);
  for (int v = 0; v < 5; ++v) var[v] = f[v * P->N + c];
}

/* ------------------------------------------------------------------------ */
/* viscous flux                                                               */
/* ------------------------------------------------------------------------ */

/* getXGrad / getYGrad / getZGrad (cartesianpatch_common.h:188-300): central
 * difference of the conserved variables at cell (i,j,k) along axis ax, one
 * sided at the patch ends (D doubled per clamped side); the y gradient is 0
 * when sizeJ() <= 2, the x and z gradients have no such guard. */
static void cell_grad(const opatch* P, const REAL* f, int ax, size_t i, size_t j, size_t k, REAL* grad) {
  const size_t n = ax == 0 ? (size_t)P->ni : (ax == 1 ? (size_t)P->nj : (size_t)P->nk);
  if (ax == 1 && !(n > 2)) {
    for (int v = 0; v < 5; ++v) grad[v] = 0;
    return;
  }
  const REAL d = ax == 0 ? P->dx : (ax == 1 ? P->dy : P->dz);
  REAL D = (REAL)0.5 / d;
  const size_t c = ax == 0 ? i : (ax == 1 ? j : k);
  size_t c1 = c - 1, c2 = c + 1;
  if (c1 >= n) {
    c1 = c;
    D *= 2;
  }
  if (c2 >= n) {
    c2 = c;
    D *= 2;
  }
  REAL var[5];
  get_var(P, f, ax == 0 ? c2 : i, ax == 1 ? c2 : j, ax == 2 ? c2 : k, grad);
  get_var(P, f, ax == 0 ? c1 : i, ax == 1 ? c1 : j, ax == 2 ? c1 : k, var);
  for (int v = 0; v < 5; ++v) {
    grad[v] -= var[v];
    grad[v] *= D;
  }
}

/* PerfectGas::mu (perfectgas.h:100-111), Sutherland's law; the float build
 * calls powf (DRNUM_SINGLE_PRECISION, drnum.h:43), the double build pow */
REAL F(gas_mu)(const REAL* var) {
  const REAL T0_suth = 300.55;
  const REAL C_suth = 111.00;
  const REAL mu0_suth = 1.781e-5;
  REAL p, T, u, v, w;
  cons_to_prim(var, &p, &T, &u, &v, &w);
#if O_IS_FLOAT
  return mu0_suth * powf(T / T0_suth, 1.5) * (T0_suth + C_suth) / (T + C_suth);
#else
  return mu0_suth * pow(T / T0_suth, 1.5) * (T0_suth + C_suth) / (T + C_suth);
#endif
}

/* derivative of the velocity component c from conserved gradients g */
static REAL dvel(const REAL* g, const REAL* var, REAL ir, int c) { return (g[c] - g[0] * var[c] * ir) * ir; }

/* gamma1/R times the derivative of T (the reference's dT_d*) */
static REAL dtemp(const REAL* g, const REAL* var, REAL ir, REAL gamma1) {
  return gamma1 * ir * ir *
         (g[4] * var[0] - g[0] * var[4] - var[1] * g[1] - var[2] * g[2] - var[3] * g[3] +
          g[0] * ir * (var[1] * var[1] + var[2] * var[2] + var[3] * var[3]));
}

/* CompressibleViscFlux::xField/yField/zField (compressibleviscflux.h:33-268)
 * for the face between the cell (i,j,k) - e_dir and (i,j,k); subtracts the
 * viscous stresses and heat flux from flux[1..4] */
void F(visc_face)(const REAL* var_l, const REAL* var_r, const REAL* gl_a, const REAL* gr_a, const REAL* gl_b,
                  const REAL* gr_b, REAL D, int dir, REAL A, REAL* flux) {
  REAL var[5], gn[5], ga[5], gb[5];
  for (int v = 0; v < 5; ++v) var[v] = (REAL)0.5 * (var_l[v] + var_r[v]);
  const REAL gamma = GAS_GAMMA;
  const REAL gamma1 = gamma - (REAL)1.0;
  const REAL ir = (REAL)1.0 / var[0];
  const REAL mu = F(gas_mu)(var);
  const REAL Pr = (REAL)0.7;
  const REAL cp = GAS_CP;
  const REAL R = GAS_R;
  const REAL k_heat = mu * cp / Pr;
  const REAL F2 = FR23 * mu, F3 = mu, F4 = k_heat / R;
  for (int v = 0; v < 5; ++v) {
    gn[v] = (var_r[v] - var_l[v]) * D;
    ga[v] = (REAL)0.5 * (gl_a[v] + gr_a[v]);
    gb[v] = (REAL)0.5 * (gl_b[v] + gr_b[v]);
  }
  /* the gradients along x, y, z: the normal one and the two tangential ones in
   * the reference's order (x face: y, z; y face: x, z; z face: x, y) */
  const REAL* gx = dir == 0 ? gn : ga;
  const REAL* gy = dir == 0 ? ga : (dir == 1 ? gn : gb);
  const REAL* gz = dir == 2 ? gn : gb;
  if (dir == 0) {
    const REAL du_dx = dvel(gx, var, ir, 1), dv_dx = dvel(gx, var, ir, 2), dw_dx = dvel(gx, var, ir, 3);
    const REAL dT_dx = dtemp(gx, var, ir, gamma1);
    const REAL du_dy = dvel(gy, var, ir, 1), dv_dy = dvel(gy, var, ir, 2);
    const REAL du_dz = dvel(gz, var, ir, 1), dw_dz = dvel(gz, var, ir, 3);
    const REAL tau_xx = F2 * ((REAL)2.0 * du_dx - dv_dy - dw_dz);
    const REAL tau_xy = F3 * (du_dy + dv_dx);
    const REAL tau_xz = F3 * (du_dz + dw_dx);
    const REAL q = F4 * dT_dx;
    flux[1] -= tau_xx * A;
    flux[2] -= tau_xy * A;
    flux[3] -= tau_xz * A;
    flux[4] -= ((tau_xx * var[1] + tau_xy * var[2] + tau_xz * var[3]) * ir + q) * A;
  } else if (dir == 1) {
    const REAL du_dx = dvel(gx, var, ir, 1), dv_dx = dvel(gx, var, ir, 2);
    const REAL du_dy = dvel(gy, var, ir, 1), dv_dy = dvel(gy, var, ir, 2), dw_dy = dvel(gy, var, ir, 3);
    const REAL dT_dy = dtemp(gy, var, ir, gamma1);
    const REAL dv_dz = dvel(gz, var, ir, 2), dw_dz = dvel(gz, var, ir, 3);
    const REAL tau_yy = F2 * (-du_dx + (REAL)2.0 * dv_dy - dw_dz);
    const REAL tau_xy = F3 * (du_dy + dv_dx);
    const REAL tau_yz = F3 * (dv_dz + dw_dy);
    const REAL q = F4 * dT_dy;
    flux[1] -= tau_xy * A;
    flux[2] -= tau_yy * A;
    flux[3] -= tau_yz * A;
    flux[4] -= ((tau_xy * var[1] + tau_yy * var[2] + tau_yz * var[3]) * ir + q) * A;
  } else {
    const REAL du_dx = dvel(gx, var, ir, 1), dw_dx = dvel(gx, var, ir, 3);
    const REAL dv_dy = dvel(gy, var, ir, 2), dw_dy = dvel(gy, var, ir, 3);
    const REAL du_dz = dvel(gz, var, ir, 1), dv_dz = dvel(gz, var, ir, 2), dw_dz = dvel(gz, var, ir, 3);
    const REAL dT_dz = dtemp(gz, var, ir, gamma1);
    const REAL tau_zz = F2 * (-du_dx - dv_dy + (REAL)2.0 * dw_dz);
    const REAL tau_xz = F3 * (du_dz + dw_dx);
    const REAL tau_yz = F3 * (dv_dz + dw_dy);
    const REAL q = F4 * dT_dz;
    flux[1] -= tau_xz * A;
    flux[2] -= tau_yz * A;
    flux[3] -= tau_zz * A;
    flux[4] -= ((tau_xz * var[1] + tau_yz * var[2] + tau_zz * var[3]) * ir + q) * A;
  }
}

/* the viscous part of the face between (i,j,k) - e_dir and (i,j,k) on field f */
static void visc_call(const opatch* P, const REAL* f, int dir, size_t i, size_t j, size_t k, REAL A, REAL* flux) {
  const size_t il = dir == 0 ? i - 1 : i, jl = dir == 1 ? j - 1 : j, kl = dir == 2 ? k - 1 : k;
  REAL vl[5], vr[5], gla[5], gra[5], glb[5], grb[5];
  get_var(P, f, il, jl, kl, vl);
  get_var(P, f, i, j, k, vr);
  const int ta = dir == 0 ? 1 : 0, tb = dir == 2 ? 1 : 2;
  /* compressibleviscflux.h:69-72 order: left a, left b, right a, right b */
  cell_grad(P, f, ta, il, jl, kl, gla);
  cell_grad(P, f, tb, il, jl, kl, glb);
  cell_grad(P, f, ta, i, j, k, gra);
  cell_grad(P, f, tb, i, j, k, grb);
  const REAL d = dir == 0 ? P->dx : (dir == 1 ? P->dy : P->dz);
  F(visc_face)(vl, vr, gla, gra, glb, grb, (REAL)1.0 / d, dir, A, flux);
}

/* one wall face call; (ta, tb) are the accumulated tangential coordinates of
 * the wall loops in cartesianiterator.h:190-308, (la, lb) the tangential indices */
static void wall_call(const opatch* P, const REAL* f0, int face, size_t i, size_t j, size_t k, REAL ta, REAL tb,
                      int64_t la, int64_t lb, REAL A, REAL* flux) {
  int dir = face / 2, plus = face % 2;
  REAL var[5];
  /* the functors read the cell inside the face: index-1 for the plus walls */
  size_t ci = i, cj = j, ck = k;
  if (plus) {
    if (dir == 0) ci = i - 1;
    if (dir == 1) cj = j - 1;
    if (dir == 2) ck = k - 1;
  }
  get_var(P, f0, ci, cj, ck, var);
  switch (P->bc[face]) {
    case O_BC_SLIP: wall_slip(var, dir, A, flux); break;
    case O_BC_TRANSMISSIVE: wall_transmissive(var, dir, A, flux); break;
    case O_BC_FARFIELD: wall_farfield(P->far_var[face], var, dir, plus, A, flux); break;
    case O_BC_JET: wall_jet(P->bc_par[face], var, dir, plus, ta, tb, la, lb, A, flux); break;
    default: break;
  }
}

/* cartesianiterator.h:87-325 for one patch (single thread order).  The main
 * block visits cells in i,j,k order and scatters each minus-face flux into the
 * two cells; the walls follow; then f0 = f1 + factor/dV * Res. */
static void patch_residual(const GRID* g, const opatch* P, REAL* res) {
  const size_t N = P->N;
  const REAL* f0 = fieldp(P, 0);
  const size_t NI = P->ni, NJ = P->nj, NK = P->nk;
  REAL Ax = P->dy * P->dz;
  REAL Ay = P->dx * P->dz;
  REAL Az = P->dx * P->dy;
  REAL flux[5], vl[5], vr[5];
  for (size_t l = 0; l < 5 * N; ++l) res[l] = 0;

  for (size_t i = 0; i < NI; ++i) {
    for (size_t j = 0; j < NJ; ++j) {
      for (size_t k = 0; k < NK; ++k) {
        size_t c = IDX(P, i, j, k);
        if (i > 0 && NI > 2) {
          memset(flux, 0, sizeof(flux));
          project(P, f0, g->recon, g->lim, i - 1, j, k, i, j, k, vl);
          project(P, f0, g->recon, g->lim, i, j, k, i - 1, j, k, vr);
          F(ausm_flux)(vl, vr, 0, Ax, flux);
          if (g->visc) visc_call(P, f0, 0, i, j, k, Ax, flux);
          size_t cm = IDX(P, i - 1, j, k);
          for (int v = 0; v < 5; ++v) {
            res[v * N + cm] -= flux[v];
            res[v * N + c] += flux[v];
          }
        }
        if (j > 0 && NJ > 2) {
          memset(flux, 0, sizeof(flux));
          project(P, f0, g->recon, g->lim, i, j - 1, k, i, j, k, vl);
          project(P, f0, g->recon, g->lim, i, j, k, i, j - 1, k, vr);
          F(ausm_flux)(vl, vr, 1, Ay, flux);
          if (g->visc) visc_call(P, f0, 1, i, j, k, Ay, flux);
          size_t cm = IDX(P, i, j - 1, k);
          for (int v = 0; v < 5; ++v) {
            res[v * N + cm] -= flux[v];
            res[v * N + c] += flux[v];
          }
        }
        if (k > 0 && NK > 2) {
          memset(flux, 0, sizeof(flux));
          project(P, f0, g->recon, g->lim, i, j, k - 1, i, j, k, vl);
          project(P, f0, g->recon, g->lim, i, j, k, i, j, k - 1, vr);
          F(ausm_flux)(vl, vr, 2, Az, flux);
          if (g->visc) visc_call(P, f0, 2, i, j, k, Az, flux);
          size_t cm = IDX(P, i, j, k - 1);
          for (int v = 0; v < 5; ++v) {
            res[v * N + cm] -= flux[v];
            res[v * N + c] += flux[v];
          }
        }
      }
    }
  }

  /* x walls (cartesianiterator.h:190-228) */
  if (NI > 2) {
    REAL y = 0.5 * P->dy;
    for (size_t j = 0; j < NJ; ++j) {
      REAL z = 0.5 * P->dz;
      for (size_t k = 0; k < NK; ++k) {
        memset(flux, 0, sizeof(flux));
        wall_call(P, f0, 0, 0, j, k, y, z, (int64_t)j, (int64_t)k, Ax, flux);
        for (int v = 0; v < 5; ++v) res[v * N + IDX(P, 0, j, k)] += flux[v];
        z += P->dz;
      }
      y += P->dy;
    }
    y = 0.5 * P->dy;
    for (size_t j = 0; j < NJ; ++j) {
      REAL z = 0.5 * P->dz;
      for (size_t k = 0; k < NK; ++k) {
        memset(flux, 0, sizeof(flux));
        wall_call(P, f0, 1, NI, j, k, y, z, (int64_t)j, (int64_t)k, Ax, flux);
        for (int v = 0; v < 5; ++v) res[v * N + IDX(P, NI - 1, j, k)] -= flux[v];
        z += P->dz;
      }
      y += P->dy;
    }
  }
  /* y walls (cartesianiterator.h:230-268) */
  if (NJ > 2) {
    REAL x = 0.5 * P->dx;
    for (size_t i = 0; i < NI; ++i) {
      REAL z = 0.5 * P->dz;
      for (size_t k = 0; k < NK; ++k) {
        memset(flux, 0, sizeof(flux));
        wall_call(P, f0, 2, i, 0, k, x, z, (int64_t)i, (int64_t)k, Ay, flux);
        for (int v = 0; v < 5; ++v) res[v * N + IDX(P, i, 0, k)] += flux[v];
        z += P->dz;
      }
      x += P->dx;
    }
    x = 0.5 * P->dx;
    for (size_t i = 0; i < NI; ++i) {
      REAL z = 0.5 * P->dz;
      for (size_t k = 0; k < NK; ++k) {
        memset(flux, 0, sizeof(flux));
        wall_call(P, f0, 3, i, NJ, k, x, z, (int64_t)i, (int64_t)k, Ay, flux);
        for (int v = 0; v < 5; ++v) res[v * N + IDX(P, i, NJ - 1, k)] -= flux[v];
        z += P->dz;
      }
      x += P->dx;
    }
  }
  /* z walls (cartesianiterator.h:270-308) */
  if (NK > 2) {
    REAL x = 0.5 * P->dx;
    for (size_t i = 0; i < NI; ++i) {
      REAL y = 0.5 * P->dy;
      for (size_t j = 0; j < NJ; ++j) {
        memset(flux, 0, sizeof(flux));
        wall_call(P, f0, 4, i, j, 0, x, y, (int64_t)i, (int64_t)j, Az, flux);
        for (int v = 0; v < 5; ++v) res[v * N + IDX(P, i, j, 0)] += flux[v];
        y += P->dy;
      }
      x += P->dx;
    }
    x = 0.5 * P->dx;
    for (size_t i = 0; i < NI; ++i) {
      REAL y = 0.5 * P->dy;
      for (size_t j = 0; j < NJ; ++j) {
        memset(flux, 0, sizeof(flux));
        wall_call(P, f0, 5, i, j, NK, x, y, (int64_t)i, (int64_t)j, Az, flux);
        for (int v = 0; v < 5; ++v) res[v * N + IDX(P, i, j, NK - 1)] -= flux[v];
        y += P->dy;
      }
      x += P->dx;
    }
  }
}

static void patch_compute(const GRID* g, opatch* P, REAL factor, REAL* res) {
  patch_residual(g, P, res);
  /* cartesianiterator.h:310-322 */
  REAL dV = P->dx * P->dy * P->dz;
  REAL patch_factor = factor / dV;
  REAL* f0 = fieldp(P, 0);
  const REAL* f1 = fieldp(P, 1);
  if (!P->flags) {
    for (size_t l = 0; l < 5 * P->N; ++l) f0[l] = f1[l] + patch_factor * res[l];
    return;
  }
  /* inactive cells keep field 0 (Patch::isActive, cartesianiterator.h:315) */
  for (size_t c = 0; c < P->N; ++c)
    if (!(P->flags[c] & O_CELL_INACTIVE))
      for (int v = 0; v < 5; ++v) f0[v * P->N + c] = f1[v * P->N + c] + patch_factor * res[v * P->N + c];
}

void F(grid_set_cell_flags)(GRID* g, int p, const unsigned char* flags) {
  opatch* P = &g->p[p];
  free(P->flags);
  P->flags = NULL;
  if (!flags) return;
  P->flags = (unsigned char*)malloc(P->N);
  memcpy(P->flags, flags, P->N);
}

/* ------------------------------------------------------------------------ */
/* coordinate transforms                                                      */
/* ------------------------------------------------------------------------ */

/* CoordTransform::CoordTransform (coordtransform.cpp:23-28) */
static void ct_identity(ctrans_t* t) {
  memset(t, 0, sizeof(*t));
  t->a[0][0] = t->a[1][1] = t->a[2][2] = 1;
}

/* MathVector::abs / normalise (mathvector_methods.h:122-127, 178-183) */
static REAL v3_abs(const REAL* v) {
  REAL l = 0;
  for (int i = 0; i < 3; ++i) l += v[i] * v[i];
  return RSQRT(l);
}
static void v3_normalise(REAL* v) {
  REAL l = v3_abs(v);
  for (int i = 0; i < 3; ++i) v[i] /= l;
}

/* CoordTransform::setMatrixFromBaseIJ (coordtransform.cpp:37-58): the columns
 * are the unit base vectors i, j (orthogonalised against i) and k = i x j */
static void ct_set_base_ij(ctrans_t* t, const REAL* bi0, const REAL* bj0) {
  REAL bi[3] = {bi0[0], bi0[1], bi0[2]}, bj[3] = {bj0[0], bj0[1], bj0[2]}, bk[3];
  v3_normalise(bi);
  v3_normalise(bj);
  /* mathvector_methods.h:86-93 */
  bk[0] = bi[1] * bj[2] - bi[2] * bj[1];
  bk[1] = bi[2] * bj[0] - bi[0] * bj[2];
  bk[2] = bi[0] * bj[1] - bi[1] * bj[0];
  REAL bk_len = v3_abs(bk);
  /* scalar product (mathvector_operators.h:208-216) */
  REAL scal_ij = 0;
  for (int i = 0; i < 3; ++i) scal_ij += bi[i] * bj[i];
  for (int i = 0; i < 3; ++i) bj[i] = bj[i] - scal_ij * bi[i];
  v3_normalise(bj);
  /* (1./len) is a double expression converted to the `real` factor of
   * operator*(const real&, MathVector) (mathvector_operators.h:184-190) */
  REAL inv_len = (REAL)(1. / bk_len);
  for (int i = 0; i < 3; ++i) bk[i] = inv_len * bk[i];
  for (int i = 0; i < 3; ++i) {
    t->a[i][0] = bi[i];
    t->a[i][1] = bj[i];
    t->a[i][2] = bk[i];
  }
}

/* CoordTransform::inverse (coordtransform.cpp:99-114) with the closed-form 3x3
 * inverse of InvSmallSquareMatrix<T,3> (smallsquarematrix.h:470-489, the
 * unsafe branch: a copied matrix has prec_safe = false) and M*v
 * (smallsquarematrix.h:209-218) for the shift */
static ctrans_t ct_inverse(const ctrans_t* c) {
  const REAL(*M)[3] = c->a;
  REAL t4 = M[0][0] * M[1][1];
  REAL t6 = M[0][0] * M[1][2];
  REAL t8 = M[0][1] * M[1][0];
  REAL t10 = M[0][2] * M[1][0];
  REAL t12 = M[0][1] * M[2][0];
  REAL t14 = M[0][2] * M[2][0];
  REAL t17 = 1 / (t4 * M[2][2] - t6 * M[2][1] - t8 * M[2][2] + t10 * M[2][1] + t12 * M[1][2] - t14 * M[1][1]);
  ctrans_t t;
  t.a[0][0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) * t17;
  t.a[0][1] = -(M[0][1] * M[2][2] - M[0][2] * M[2][1]) * t17;
  t.a[0][2] = -(-M[0][1] * M[1][2] + M[0][2] * M[1][1]) * t17;
  t.a[1][0] = -(M[1][0] * M[2][2] - M[1][2] * M[2][0]) * t17;
  t.a[1][1] = (M[0][0] * M[2][2] - t14) * t17;
  t.a[1][2] = -(t6 - t10) * t17;
  t.a[2][0] = -(-M[1][0] * M[2][1] + M[1][1] * M[2][0]) * t17;
  t.a[2][1] = -(M[0][0] * M[2][1] - t12) * t17;
  t.a[2][2] = (t4 - t8) * t17;
  const REAL v[3] = {-c->b[0], -c->b[1], -c->b[2]};
  for (int i = 0; i < 3; ++i) {
    t.b[i] = 0;
    for (int j = 0; j < 3; ++j) t.b[i] += t.a[i][j] * v[j];
  }
  return t;
}

/* CoordTransform::concatenate (coordtransform.h:136-165): ret = t * c, i.e.
 * ret(x) = t(c(x)) */
static ctrans_t ct_concatenate(const ctrans_t* t, const ctrans_t* c) {
  ctrans_t r;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) r.a[i][j] = t->a[i][0] * c->a[0][j] + t->a[i][1] * c->a[1][j] + t->a[i][2] * c->a[2][j];
    r.b[i] = t->a[i][0] * c->b[0] + t->a[i][1] * c->b[1] + t->a[i][2] * c->b[2] + t->b[i];
  }
  return r;
}

/* CoordTransform::transform (coordtransform.h:105-111) */
static void ct_transform(const ctrans_t* t, REAL x, REAL y, REAL z, REAL* out) {
  for (int i = 0; i < 3; ++i) out[i] = t->a[i][0] * x + t->a[i][1] * y + t->a[i][2] * z + t->b[i];
}

/* Patch::setupTransformation (patch.cpp:265-278): this2inertial = (base matrix,
 * origin); m_TransformInertial2This = its inverse, whose stored inverse
 * (CoordTransformVV::computeInverse, coordtransformvv.cpp:105-108) is
 * inverted once more */
static void setup_transformation(const REAL* o, const REAL* bi, const REAL* bj, ctrans_t* i2t, ctrans_t* t2i) {
  ctrans_t this2inertial;
  ct_identity(&this2inertial);
  for (int a = 0; a < 3; ++a) this2inertial.b[a] = o[a];
  ct_set_base_ij(&this2inertial, bi, bj);
  *i2t = ct_inverse(&this2inertial);
  *t2i = ct_inverse(i2t);
}

/* ------------------------------------------------------------------------ */
/* grid set-up                                                                */
/* ------------------------------------------------------------------------ */

/* cartesianpatch.cpp:164-223 computeDeltas, :314-390 buildRegions (no extra protection layers) */
static void patch_setup(opatch* P, const o_patch_desc* d) {
  memset(P, 0, sizeof(*P));
  P->ni = d->ni;
  P->nj = d->nj;
  P->nk = d->nk;
  P->njk = (size_t)d->nj * d->nk;
  P->N = (size_t)d->ni * P->njk;
  for (int a = 0; a < 6; ++a) P->seek[a] = d->seek[a];
  for (int a = 0; a < 3; ++a) {
    P->o[a] = (REAL)d->origin[a];
    P->L[a] = (REAL)d->length[a];
    P->base_i[a] = (REAL)d->base_i[a];
    P->base_j[a] = (REAL)d->base_j[a];
  }
  setup_transformation(P->o, P->base_i, P->base_j, &P->i2t, &P->t2i);
  P->dx = P->L[0] / (REAL)(size_t)P->ni;
  P->dy = P->L[1] / (REAL)(size_t)P->nj;
  P->dz = P->L[2] / (REAL)(size_t)P->nk;
  P->xccmin = 0.5 * P->dx;
  P->yccmin = 0.5 * P->dy;
  P->zccmin = 0.5 * P->dz;
  int n[3] = {P->ni, P->nj, P->nk};
  for (int a = 0; a < 3; ++a) {
    P->dzf[a] = (size_t)P->seek[2 * a];
    P->dza[a] = (size_t)n[a] - (size_t)P->seek[2 * a + 1];
  }
  for (int f = 0; f < 6; ++f) {
    P->bc[f] = d->bc[f];
    memcpy(P->bc_par[f], d->bc_par[f], sizeof(P->bc_par[f]));
    if (d->bc[f] == O_BC_FARFIELD) {
      F(prim_to_cons)((REAL)d->bc_par[f][0], (REAL)d->bc_par[f][1], (REAL)d->bc_par[f][2],
                      (REAL)d->bc_par[f][3], (REAL)d->bc_par[f][4], P->far_var[f]);
    }
  }
  P->data = (REAL*)calloc(3 * 5 * P->N, sizeof(REAL));
}

GRID* F(grid_create)(int n_patches, const o_patch_desc* pd, int recon, int lim) {
  GRID* g = (GRID*)calloc(1, sizeof(GRID));
  g->np = n_patches;
  g->recon = recon;
  g->lim = lim;
  g->nthreads = 1;
  g->p = (opatch*)calloc((size_t)n_patches, sizeof(opatch));
  for (int p = 0; p < n_patches; ++p) patch_setup(&g->p[p], &pd[p]);
  return g;
}

static void free_deps(opatch* P) {
  free(P->recv);
  free(P->donors);
  free(P->rec_concat);
  free(P->donor_idx);
  free(P->donor_w);
  free(P->recv64);
  free(P->rec_concat64);
  free(P->donor_idx64);
  P->recv = NULL;
  P->donors = NULL;
  P->rec_concat = NULL;
  P->donor_idx = NULL;
  P->donor_w = NULL;
  P->recv64 = P->rec_concat64 = P->donor_idx64 = NULL;
  P->n_recv = P->n_rec_concat = P->n_wi = 0;
  P->n_donors = 0;
}

static void free_ls(opatch* P) {
  if (P->inner) {
    for (int l = 0; l < P->n_inner; ++l) free(P->inner[l].v);
    free(P->inner);
  }
  if (P->outer) {
    for (int l = 0; l < P->n_outer; ++l) free(P->outer[l].v);
    free(P->outer);
  }
  free(P->G);
  P->inner = P->outer = NULL;
  P->G = NULL;
  P->n_inner = P->n_outer = 0;
}

void F(grid_destroy)(GRID* g) {
  if (!g) return;
  for (int p = 0; p < g->np; ++p) {
    free(g->p[p].data);
    free(g->p[p].flags);
    free_deps(&g->p[p]);
    free_ls(&g->p[p]);
  }
  free(g->p);
  free(g);
}

int F(grid_num_patches)(const GRID* g) { return g->np; }
size_t F(grid_patch_cells)(const GRID* g, int p) { return g->p[p].N; }
REAL* F(grid_field)(GRID* g, int p, int field) { return fieldp(&g->p[p], field); }
REAL F(grid_dx)(const GRID* g, int p, int axis) {
  return axis == 0 ? g->p[p].dx : (axis == 1 ? g->p[p].dy : g->p[p].dz);
}
void F(grid_set_threads)(GRID* g, int n) { g->nthreads = n < 1 ? 1 : n; }
void F(grid_set_viscous)(GRID* g, int on) { g->visc = on != 0; }

/* patchgrid.cpp:645-651 */
void F(grid_set_const)(GRID* g, int field, const REAL* var) {
  for (int p = 0; p < g->np; ++p) {
    REAL* f = fieldp(&g->p[p], field);
    for (int v = 0; v < 5; ++v)
      for (size_t l = 0; l < g->p[p].N; ++l) f[v * g->p[p].N + l] = var[v];
  }
}

void F(grid_copy_field)(GRID* g, int src, int dst) {
  for (int p = 0; p < g->np; ++p)
    memcpy(fieldp(&g->p[p], dst), fieldp(&g->p[p], src), 5 * g->p[p].N * sizeof(REAL));
}

/* PatchIterator::computeAll -> CartesianIterator::compute, patch by patch.
 * Patches are independent, so they may run on separate threads without
 * changing any result (each patch's residual is summed in one thread). */
void F(grid_compute)(GRID* g, REAL factor) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(g->nthreads)
#endif
  for (int p = 0; p < g->np; ++p) {
    REAL* res = (REAL*)malloc(5 * g->p[p].N * sizeof(REAL));
    patch_compute(g, &g->p[p], factor, res);
    free(res);
  }
}

void F(grid_residual)(GRID* g, int p, REAL* res) { patch_residual(g, &g->p[p], res); }

/* ------------------------------------------------------------------------ */
/* donor tables                                                               */
/* ------------------------------------------------------------------------ */

/* weightedset.h:374-404 Unify: sort (index, weight) pairs, add weights of equal indices */
static void ws_unify(wset_t* s) {
  for (size_t a = 1; a < s->n; ++a) {
    size_t ti = s->idx[a];
    REAL tw = s->w[a];
    size_t b = a;
    while (b > 0 && (s->idx[b - 1] > ti || (s->idx[b - 1] == ti && s->w[b - 1] > tw))) {
      s->idx[b] = s->idx[b - 1];
      s->w[b] = s->w[b - 1];
      --b;
    }
    s->idx[b] = ti;
    s->w[b] = tw;
  }
  if (s->n == 0) return;
  size_t hold = 0;
  for (size_t it = 1; it < s->n; ++it) {
    if (s->idx[it] == s->idx[hold]) {
      s->w[hold] += s->w[it];
    } else {
      ++hold;
      s->idx[hold] = s->idx[it];
      s->w[hold] = s->w[it];
    }
  }
  s->n = hold + 1;
}

/* weightedset.h:569-602 (relative = false, keep_sum = false) */
static void ws_eliminate_below(wset_t* s, REAL eps) {
  ws_unify(s);
  size_t m = 0;
  for (size_t a = 0; a < s->n; ++a) {
    REAL t = s->w[a];
    if (t < 0) t = -t;
    if (!(t < eps)) {
      s->idx[m] = s->idx[a];
      s->w[m] = s->w[a];
      ++m;
    }
  }
  s->n = m;
}

/* weightedset.h:605-619 */
static void ws_adjust_sum_shift(wset_t* s, REAL weight_sum) {
  if (s->n == 0) return;
  REAL old = 0;
  for (size_t a = 0; a < s->n; ++a) old += s->w[a];
  REAL corr = (weight_sum - old) / (REAL)s->n;
  for (size_t a = 0; a < s->n; ++a) s->w[a] += corr;
}

/* cartesianpatch.cpp:1053-1162 linearCCInterpolCoeffs.  Returns 1 inside,
 * 0 outside, -1 on the reference's BUG conditions. */
static int lin_coeff(REAL s_test, size_t first, size_t afterlast, REAL ds, size_t num, size_t* ref, REAL* coeff) {
  const REAL m_eps = (REAL)1.e-5; /* cartesianpatch.cpp:40 */
  REAL rs_test = s_test / ds;
  REAL rs0 = (float)(first);
  REAL rs1 = rs0 + 0.5;
  REAL rs2 = (float)(afterlast)-0.5;
  REAL rs3 = rs2 + 0.5;
  int sector;
  if (rs_test < rs0) return 0;
  else if (rs_test < rs1) sector = 1;
  else if (rs_test < rs2) sector = 2;
  else if (rs_test < rs3) sector = 3;
  else return 0;
  if (sector == 1) {
    *ref = first;
    *coeff = 0.;
  } else if (sector == 2) {
    long l = (long)(rs_test - 0.5 + (float)(num)) - (long)num;
    if ((size_t)l == first - 1) {
      l = (long)first;
    } else if ((size_t)l == afterlast - 1) {
      l = (long)afterlast - 2;
    } else if (l < -1 || (size_t)l > afterlast - 1) {
      return -1;
    }
    *ref = (size_t)l;
    REAL rs_cell_ref = 0.5 + (float)(*ref);
    *coeff = rs_test - rs_cell_ref;
  } else {
    *ref = afterlast - 2;
    *coeff = 1.;
  }
  if (*ref < first || *ref > afterlast - 2) return -1;
  if (*coeff < -m_eps) return -1;
  if (*coeff > (1. + m_eps)) return -1;
  return 1;
}

/* cartesianpatch.cpp:1166-1299 computeCCDataInterpolCoeffs_V1 */
static int cc_interp(const opatch* Q, REAL x, REAL y, REAL z, wset_t* ws) {
  size_t ic, jc, kc;
  REAL xc, yc, zc;
  int r;
  if ((r = lin_coeff(x, Q->dzf[0], Q->dza[0], Q->dx, (size_t)Q->ni, &ic, &xc)) != 1) return r;
  if ((r = lin_coeff(y, Q->dzf[1], Q->dza[1], Q->dy, (size_t)Q->nj, &jc, &yc)) != 1) return r;
  if ((r = lin_coeff(z, Q->dzf[2], Q->dza[2], Q->dz, (size_t)Q->nk, &kc, &zc)) != 1) return r;
  REAL w[8];
  w[0] = (1. - xc) * (1. - yc) * (1. - zc);
  w[1] = (1. - xc) * (1. - yc) * zc;
  w[2] = (1. - xc) * yc * (1. - zc);
  w[3] = (1. - xc) * yc * zc;
  w[4] = xc * (1. - yc) * (1. - zc);
  w[5] = xc * (1. - yc) * zc;
  w[6] = xc * yc * (1. - zc);
  w[7] = xc * yc * zc;
  size_t ic1 = ic + 1, jc1 = jc + 1, kc1 = kc + 1;
  if (ic1 + 1 > (size_t)Q->ni) ic1 = ic;
  if (jc1 + 1 > (size_t)Q->nj) jc1 = jc;
  if (kc1 + 1 > (size_t)Q->nk) kc1 = kc;
  ws->n = 8;
  ws->idx[0] = IDX(Q, ic, jc, kc);
  ws->idx[1] = IDX(Q, ic, jc, kc1);
  ws->idx[2] = IDX(Q, ic, jc1, kc);
  ws->idx[3] = IDX(Q, ic, jc1, kc1);
  ws->idx[4] = IDX(Q, ic1, jc, kc);
  ws->idx[5] = IDX(Q, ic1, jc, kc1);
  ws->idx[6] = IDX(Q, ic1, jc1, kc);
  ws->idx[7] = IDX(Q, ic1, jc1, kc1);
  for (int a = 0; a < 8; ++a) ws->w[a] = w[a];
  const REAL m_eps = (REAL)1.e-5;
  ws_eliminate_below(ws, 10 * m_eps);
  ws_adjust_sum_shift(ws, (REAL)1.);
  return 1;
}

typedef struct {
  size_t ll, l;
  wset_t ws;
} wsentry_t;

typedef struct {
  int patch;
  size_t n, cap;
  wsentry_t* v;
  REAL a[3][3]; /* turn of vector variables, donor -> receiver */
} neigh_t;

/* patch.cpp:264-317 insertNeighbour + cartesianpatch.cpp:494-526 computeDependencies.
 * The relative transform is CoordTransformVV::setTransFromTo(receiver, donor)
 * (coordtransformvv.cpp:65-71): m_transform = (receiver this->inertial)
 * .concatenate(donor inertial->this), applied to the receiver's cell centres,
 * and m_transform_inverse = (donor this->inertial).concatenate(receiver
 * inertial->this), whose matrix turns the donor's vector variables
 * (InterCoeffWS::setCoordTransform(extractReverse()), patch.cpp:293-294, 178-186).
 * concatenate(c) applies c first (coordtransform.h:136-165), so the
 * reference's composition equals the geometric one (donor inertial->this after
 * receiver this->inertial) when the two commute, e.g. when one of the two
 * patches is unrotated with its origin at the inertial origin, or when both
 * are axis aligned (a translation by o_receiver - o_donor). */
static int collect_neighbour(const opatch* P, const opatch* Q, const REAL* oq, size_t* hits, neigh_t* nb) {
  ctrans_t q_i2t, q_t2i;
  setup_transformation(oq, Q->base_i, Q->base_j, &q_i2t, &q_t2i);
  const ctrans_t rel = ct_concatenate(&P->t2i, &q_i2t);
  const ctrans_t turn = ct_concatenate(&q_t2i, &P->i2t);
  memcpy(nb->a, turn.a, sizeof(nb->a));
  for (size_t ll = 0; ll < P->n_recv; ++ll) {
    size_t l = P->recv[ll];
    size_t i = l / P->njk, rest = l - i * P->njk, j = rest / (size_t)P->nk, k = rest - j * (size_t)P->nk;
    REAL x = P->xccmin + (REAL)i * P->dx;
    REAL y = P->yccmin + (REAL)j * P->dy;
    REAL z = P->zccmin + (REAL)k * P->dz;
    REAL xt[3];
    ct_transform(&rel, x, y, z, xt);
    wset_t ws;
    int r = cc_interp(Q, xt[0], xt[1], xt[2], &ws);
    if (r < 0) return -1;
    if (r == 1) {
      if (nb->n == nb->cap) {
        nb->cap = nb->cap ? 2 * nb->cap : 256;
        nb->v = (wsentry_t*)realloc(nb->v, nb->cap * sizeof(wsentry_t));
      }
      nb->v[nb->n].ll = ll;
      nb->v[nb->n].l = l;
      nb->v[nb->n].ws = ws;
      nb->n++;
      hits[ll]++;
    }
  }
  return 0;
}

/* PatchGrid::computeDependencies (patchgrid.cpp:119-259).  Candidate donors are
 * all other patches in ascending order (the hash raster of the reference only
 * prunes pairs that can never interpolate), followed -- for periodic axes, an
 * extension not in the reference -- by periodic images in the shift order
 * (-1,0,1)^3.  `periodic` holds the domain period per axis (0 = not periodic). */
int F(grid_compute_dependencies)(GRID* g, const double* periodic) {
  /* extractSeekCells + compactReceiveCellLists (cartesianpatch.cpp:393-491, patch.cpp:468-486) */
  for (int p = 0; p < g->np; ++p) {
    opatch* P = &g->p[p];
    free_deps(P);
    P->recv = (size_t*)malloc(P->N * sizeof(size_t));
    P->n_recv = 0;
    for (size_t i = 0; i < (size_t)P->ni; ++i)
      for (size_t j = 0; j < (size_t)P->nj; ++j)
        for (size_t k = 0; k < (size_t)P->nk; ++k) {
          int s = (i < (size_t)P->seek[0]) || (i >= (size_t)P->ni - P->seek[1]) || (j < (size_t)P->seek[2]) ||
                  (j >= (size_t)P->nj - P->seek[3]) || (k < (size_t)P->seek[4]) ||
                  (k >= (size_t)P->nk - P->seek[5]);
          if (s) P->recv[P->n_recv++] = IDX(P, i, j, k);
        }
  }
  int err = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(g->nthreads) reduction(| : err)
#endif
  for (int p = 0; p < g->np; ++p) {
    opatch* P = &g->p[p];
    if (P->n_recv == 0) continue;
    size_t* hits = (size_t*)calloc(P->n_recv, sizeof(size_t));
    int n_nb = 0, cap_nb = 0;
    neigh_t* nbs = NULL;
    int shifts[27][3];
    int n_shifts = 0;
    for (int a = -1; a <= 1; ++a)
      for (int b = -1; b <= 1; ++b)
        for (int c = -1; c <= 1; ++c) {
          if ((a != 0 && !(periodic && periodic[0] > 0)) || (b != 0 && !(periodic && periodic[1] > 0)) ||
              (c != 0 && !(periodic && periodic[2] > 0)))
            continue;
          if (a == 0 && b == 0 && c == 0) continue;
          shifts[n_shifts][0] = a;
          shifts[n_shifts][1] = b;
          shifts[n_shifts][2] = c;
          n_shifts++;
        }
    for (int q = 0; q < g->np && !err; ++q) {
      for (int s = -1; s < n_shifts; ++s) {
        if (s < 0 && q == p) continue;
        const opatch* Q = &g->p[q];
        REAL oq[3];
        for (int a = 0; a < 3; ++a) {
          if (s < 0 || shifts[s][a] == 0)
            oq[a] = Q->o[a];
          else
            oq[a] = Q->o[a] + (REAL)(shifts[s][a] * periodic[a]);
        }
        if (n_nb == cap_nb) {
          cap_nb = cap_nb ? 2 * cap_nb : 8;
          nbs = (neigh_t*)realloc(nbs, (size_t)cap_nb * sizeof(neigh_t));
        }
        neigh_t* nb = &nbs[n_nb];
        memset(nb, 0, sizeof(*nb));
        nb->patch = q;
        if (collect_neighbour(P, Q, oq, hits, nb) < 0) {
          err = 1;
          free(nb->v);
          break;
        }
        if (nb->n > 0)
          n_nb++;
        else
          free(nb->v);
      }
    }
    if (!err) {
      /* Patch::finalizeDependencies (patch.cpp:319-465) */
      size_t* new_from_old = (size_t*)malloc(P->n_recv * sizeof(size_t));
      size_t shift = 0;
      for (size_t ll = 0; ll < P->n_recv; ++ll) {
        new_from_old[ll] = ll - shift;
        if (hits[ll] == 0) shift++;
      }
      P->n_no_donor = shift;
      size_t m = 0;
      for (size_t ll = 0; ll < P->n_recv; ++ll) {
        if (hits[ll] > 0) {
          P->recv[m] = P->recv[ll];
          hits[m] = hits[ll];
          ++m;
        }
      }
      P->n_recv = m;
      /* reindexIndReceive + adjust2Average (intercoeffws.h:224-242) */
      for (int d = 0; d < n_nb; ++d)
        for (size_t e = 0; e < nbs[d].n; ++e) {
          size_t ind = new_from_old[nbs[d].v[e].ll];
          nbs[d].v[e].ll = ind;
          if (hits[ind] > 1) {
            REAL sc = (REAL)(1. / hits[ind]);
            for (size_t a = 0; a < nbs[d].v[e].ws.n; ++a) nbs[d].v[e].ws.w[a] *= sc;
          }
        }
      free(new_from_old);
      /* InterCoeffPad::build (intercoeffpad.h:139-170) + buildDonorTransferData (patch.cpp:128-218) */
      P->n_donors = n_nb;
      P->donors = (donor_t_*)calloc((size_t)(n_nb > 0 ? n_nb : 1), sizeof(donor_t_));
      size_t nrc = 0, nwi = 0;
      for (int d = 0; d < n_nb; ++d) {
        size_t stride = 0;
        for (size_t e = 0; e < nbs[d].n; ++e)
          if (nbs[d].v[e].ws.n > stride) stride = nbs[d].v[e].ws.n;
        P->donors[d].patch = nbs[d].patch;
        P->donors[d].nrec = nbs[d].n;
        P->donors[d].stride = stride;
        P->donors[d].rec_start = nrc;
        P->donors[d].wi_start = nwi;
        memcpy(P->donors[d].a, nbs[d].a, sizeof(P->donors[d].a));
        nrc += nbs[d].n;
        nwi += nbs[d].n * stride;
      }
      P->n_rec_concat = nrc;
      P->n_wi = nwi;
      P->rec_concat = (size_t*)malloc((nrc ? nrc : 1) * sizeof(size_t));
      P->donor_idx = (size_t*)malloc((nwi ? nwi : 1) * sizeof(size_t));
      P->donor_w = (REAL*)malloc((nwi ? nwi : 1) * sizeof(REAL));
      for (int d = 0; d < n_nb; ++d) {
        size_t stride = P->donors[d].stride;
        for (size_t e = 0; e < nbs[d].n; ++e) {
          P->rec_concat[P->donors[d].rec_start + e] = nbs[d].v[e].l;
          size_t start = P->donors[d].wi_start + e * stride;
          const wset_t* ws = &nbs[d].v[e].ws;
          for (size_t a = 0; a < ws->n; ++a) {
            P->donor_idx[start + a] = ws->idx[a];
            P->donor_w[start + a] = ws->w[a];
          }
          for (size_t a = ws->n; a < stride; ++a) {
            P->donor_idx[start + a] = P->donor_idx[start];
            P->donor_w[start + a] = 0.;
          }
        }
      }
      P->recv64 = (int64_t*)malloc((P->n_recv ? P->n_recv : 1) * sizeof(int64_t));
      P->rec_concat64 = (int64_t*)malloc((nrc ? nrc : 1) * sizeof(int64_t));
      P->donor_idx64 = (int64_t*)malloc((nwi ? nwi : 1) * sizeof(int64_t));
      for (size_t a = 0; a < P->n_recv; ++a) P->recv64[a] = (int64_t)P->recv[a];
      for (size_t a = 0; a < nrc; ++a) P->rec_concat64[a] = (int64_t)P->rec_concat[a];
      for (size_t a = 0; a < nwi; ++a) P->donor_idx64[a] = (int64_t)P->donor_idx[a];
    }
    for (int d = 0; d < n_nb; ++d) free(nbs[d].v);
    free(nbs);
    free(hits);
  }
  g->error = err;
  return err ? -1 : 0;
}

int F(grid_num_receivers)(const GRID* g, int p) { return (int)g->p[p].n_recv; }

void F(grid_donor_matrix)(const GRID* g, int p, int d, double* a9) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) a9[3 * r + c] = (double)g->p[p].donors[d].a[r][c];
}

void F(grid_patch_transform)(const GRID* g, int p, double* a12) {
  const ctrans_t* t = &g->p[p].t2i;
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) a12[3 * r + c] = (double)t->a[r][c];
    a12[9 + r] = (double)t->b[r];
  }
}
int F(grid_num_no_donor)(const GRID* g, int p) { return (int)g->p[p].n_no_donor; }
const int64_t* F(grid_receivers)(const GRID* g, int p) { return g->p[p].recv64; }
int F(grid_num_donors)(const GRID* g, int p) { return g->p[p].n_donors; }
void F(grid_donor_info)(const GRID* g, int p, int d, int64_t* out) {
  const donor_t_* D = &g->p[p].donors[d];
  out[0] = D->patch;
  out[1] = (int64_t)D->nrec;
  out[2] = (int64_t)D->stride;
  out[3] = (int64_t)D->rec_start;
  out[4] = (int64_t)D->wi_start;
}
int64_t F(grid_num_concat)(const GRID* g, int p, int which) {
  return which == 0 ? (int64_t)g->p[p].n_rec_concat : (int64_t)g->p[p].n_wi;
}
const int64_t* F(grid_rec_concat)(const GRID* g, int p) { return g->p[p].rec_concat64; }
const int64_t* F(grid_donor_index)(const GRID* g, int p) { return g->p[p].donor_idx64; }
const REAL* F(grid_donor_weight)(const GRID* g, int p) { return g->p[p].donor_w; }

/* patch.cpp:573-653 accessDonorDataDirect, all patches (PatchIterator::copyDonorData).
 * The vector rotation (:627-641) is the identity for axis-aligned patches. */
void F(grid_exchange)(GRID* g, int field) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(g->nthreads)
#endif
  for (int p = 0; p < g->np; ++p) {
    opatch* P = &g->p[p];
    REAL* f = fieldp(P, field);
    for (size_t ll = 0; ll < P->n_recv; ++ll)
      for (int v = 0; v < 5; ++v) f[v * P->N + P->recv[ll]] = 0;
    for (int d = 0; d < P->n_donors; ++d) {
      const donor_t_* D = &P->donors[d];
      const opatch* Q = &g->p[D->patch];
      const REAL* fd = fieldp(Q, field);
      for (size_t ll = 0; ll < D->nrec; ++ll) {
        size_t rec = P->rec_concat[D->rec_start + ll];
        size_t start = D->wi_start + ll * D->stride;
        REAL inter[5] = {0., 0., 0., 0., 0.};
        for (size_t s = 0; s < D->stride; ++s) {
          size_t dc = P->donor_idx[start + s];
          REAL w = P->donor_w[start + s];
          for (int v = 0; v < 5; ++v) inter[v] += fd[v * Q->N + dc] * w;
        }
        /* the vector variable of the compressible state, momentum at index 1
         * (PatchGrid::defineVectorVar(1), drnumJetDemo/main.h:99), turned into
         * the receiver's frame (patch.cpp:627-641) */
        const REAL(*A)[3] = D->a;
        REAL u = A[0][0] * inter[1] + A[0][1] * inter[2] + A[0][2] * inter[3];
        REAL vv = A[1][0] * inter[1] + A[1][1] * inter[2] + A[1][2] * inter[3];
        REAL ww = A[2][0] * inter[1] + A[2][1] * inter[2] + A[2][2] * inter[3];
        inter[1] = u;
        inter[2] = vv;
        inter[3] = ww;
        for (int v = 0; v < 5; ++v) f[v * P->N + rec] += inter[v];
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* level set                                                                  */
/* ------------------------------------------------------------------------ */

static void ls_push(lsvec_t* L, size_t cell, REAL g) {
  if (L->n == L->cap) {
    L->cap = L->cap ? 2 * L->cap : 256;
    L->v = (lscell_t*)realloc(L->v, L->cap * sizeof(lscell_t));
  }
  memset(&L->v[L->n], 0, sizeof(lscell_t));
  L->v[L->n].cell = cell;
  L->v[L->n].g = g;
  L->n++;
}

/* cartesianpatch.cpp:1512-1549 (order k-, k+, j-, j+, i-, i+) */
static int face_neighbours(const opatch* P, size_t l, size_t* nb) {
  size_t i = l / P->njk, rest = l - i * P->njk, j = rest / (size_t)P->nk, k = rest - j * (size_t)P->nk;
  int n = 0;
  size_t cand[6][3] = {{i, j, k - 1}, {i, j, k + 1}, {i, j - 1, k}, {i, j + 1, k}, {i - 1, j, k}, {i + 1, j, k}};
  for (int a = 0; a < 6; ++a)
    if (in_range(P, cand[a][0], cand[a][1], cand[a][2])) nb[n++] = IDX(P, cand[a][0], cand[a][1], cand[a][2]);
  return n;
}

/* cartesianpatch.cpp:1592-1657 computeNablaVar */
static void nabla(const opatch* P, const REAL* var, size_t l, REAL* g) {
  size_t i = l / P->njk, rest = l - i * P->njk, j = rest / (size_t)P->nk, k = rest - j * (size_t)P->nk;
  size_t n[3] = {(size_t)P->ni, (size_t)P->nj, (size_t)P->nk};
  REAL d[3] = {P->dx, P->dy, P->dz};
  size_t c[3] = {i, j, k};
  for (int a = 0; a < 3; ++a) {
    size_t cp[3] = {c[0], c[1], c[2]}, cm[3] = {c[0], c[1], c[2]};
    cp[a] = c[a] + 1;
    cm[a] = c[a] - 1;
    int skip_p = !(cp[a] <= n[a] - 1);
    int skip_m = !(cm[a] <= n[a] - 1);
    size_t lp = IDX(P, cp[0], cp[1], cp[2]), lm = IDX(P, cm[0], cm[1], cm[2]);
    if (!skip_p && !skip_m)
      g[a] = (var[lp] - var[lm]) / (2. * d[a]);
    else if (skip_p)
      g[a] = (var[l] - var[lm]) / d[a];
    else
      g[a] = (var[lp] - var[l]) / d[a];
  }
}

/* levelsetobject.cpp:44-80 update + :83-388 extractBCellLayers, with the level
 * set min_s(|x - c_s| - r_s) (spherelevelset.cpp:19-32, combilevelsetor.cpp:15-26). */
int F(grid_levelset)(GRID* g, int n_spheres, const double* spheres, int n_inner, int n_outer,
                     double min_inner_rel) {
  g->have_ls = 1;
  for (int p = 0; p < g->np; ++p) {
    opatch* P = &g->p[p];
    free_ls(P);
    P->G = (REAL*)malloc(P->N * sizeof(REAL));
    P->n_inner = n_inner;
    P->n_outer = n_outer;
    P->inner = (lsvec_t*)calloc((size_t)(n_inner > 0 ? n_inner : 1), sizeof(lsvec_t));
    P->outer = (lsvec_t*)calloc((size_t)(n_outer > 0 ? n_outer : 1), sizeof(lsvec_t));
    REAL* var = P->G;
    for (size_t l = 0; l < P->N; ++l) {
      size_t i = l / P->njk, rest = l - i * P->njk, j = rest / (size_t)P->nk, k = rest - j * (size_t)P->nk;
      /* Patch::xyzoCell (patch.cpp:91-98): m_TransformInertial2This.transformReverse */
      REAL xo[3];
      ct_transform(&P->t2i, P->xccmin + (REAL)i * P->dx, P->yccmin + (REAL)j * P->dy, P->zccmin + (REAL)k * P->dz,
                   xo);
      REAL xc = xo[0], yc = xo[1], zc = xo[2];
      REAL dist = 0;
      for (int s = 0; s < n_spheres; ++s) {
        REAL ddx = xc - (REAL)spheres[4 * s + 0];
        REAL ddy = yc - (REAL)spheres[4 * s + 1];
        REAL ddz = zc - (REAL)spheres[4 * s + 2];
        REAL q = ddx * ddx + ddy * ddy + ddz * ddz;
        REAL dh = RSQRT(q) - (REAL)spheres[4 * s + 3];
        if (s == 0 || dist > dh) dist = dh;
      }
      var[l] = dist;
    }
    char* marker = (char*)calloc(P->N, 1);
    REAL gmax = var[0], gmin = var[0];
    for (size_t l = 0; l < P->N; ++l) {
      if (gmin > var[l]) gmin = var[l];
      if (gmax < var[l]) gmax = var[l];
    }
    REAL g_eps = (gmax - gmin) * 1.e-7;
    size_t nb[6];
    /* 0th layers */
    for (size_t l = 0; l < P->N; ++l) {
      REAL gv = var[l];
      int n = face_neighbours(P, l, nb);
      int other = 0;
      for (int a = 0; a < n; ++a) {
        REAL t1 = (var[nb[a]] + g_eps) * (gv - g_eps);
        REAL t2 = (var[nb[a]] - g_eps) * (gv + g_eps);
        if (t1 < 0. || t2 < 0.) {
          other = 1;
          marker[l] = 1;
          break;
        }
      }
      if (other) {
        if (gv < 0.) {
          if (n_inner > 0) ls_push(&P->inner[0], l, gv);
        } else {
          if (n_outer > 0) ls_push(&P->outer[0], l, gv);
        }
      }
    }
    /* minimum relative distance replacement (levelsetobject.cpp:164-222) */
    if (min_inner_rel > 0. && n_inner > 0) {
      REAL mrel = (REAL)min_inner_rel;
      size_t ll = 0;
      while (ll < P->inner[0].n) {
        size_t l = P->inner[0].v[ll].cell;
        int n = face_neighbours(P, l, nb);
        REAL ming = var[nb[0]], maxg = var[nb[0]];
        for (int a = 1; a < n; ++a) {
          if (ming > var[nb[a]]) ming = var[nb[a]];
          if (maxg < var[nb[a]]) maxg = var[nb[a]];
        }
        REAL gdiff = 0.5 * (maxg - ming);
        REAL mind = mrel * gdiff;
        if (var[l] > -mind) {
          memmove(&P->inner[0].v[ll], &P->inner[0].v[ll + 1], (P->inner[0].n - ll - 1) * sizeof(lscell_t));
          P->inner[0].n--;
          marker[l] = 0;
          for (int a = 0; a < n; ++a) {
            size_t lc = nb[a];
            if (!marker[lc] && var[lc] < var[l]) {
              ls_push(&P->inner[0], lc, var[lc]);
              marker[lc] = 1;
            }
          }
        } else {
          ++ll;
        }
      }
    }
    /* further layers (levelsetobject.cpp:237-291) */
    for (int layer = 1; layer < n_inner; ++layer) {
      for (size_t ll = 0; ll < P->inner[layer - 1].n; ++ll) {
        int n = face_neighbours(P, P->inner[layer - 1].v[ll].cell, nb);
        for (int a = 0; a < n; ++a)
          if (!marker[nb[a]] && var[nb[a]] < 0.) {
            ls_push(&P->inner[layer], nb[a], var[nb[a]]);
            marker[nb[a]] = 1;
          }
      }
    }
    for (int layer = 1; layer < n_outer; ++layer) {
      for (size_t ll = 0; ll < P->outer[layer - 1].n; ++ll) {
        int n = face_neighbours(P, P->outer[layer - 1].v[ll].cell, nb);
        for (int a = 0; a < n; ++a)
          if (!marker[nb[a]] && var[nb[a]] > 0.) {
            ls_push(&P->outer[layer], nb[a], var[nb[a]]);
            marker[nb[a]] = 1;
          }
      }
    }
    free(marker);
    /* gradients and mirror tables (levelsetobject.cpp:294-388) */
    for (int layer = 0; layer < n_inner; ++layer) {
      lsvec_t* L = &P->inner[layer];
      for (size_t ll = 0; ll < L->n; ++ll) {
        lscell_t* c = &L->v[ll];
        REAL gg[3];
        nabla(P, var, c->cell, gg);
        REAL len = 0;
        for (int a = 0; a < 3; ++a) len += gg[a] * gg[a];
        len = RSQRT(len);
        for (int a = 0; a < 3; ++a) gg[a] /= len;
        c->gx = gg[0];
        c->gy = gg[1];
        c->gz = gg[2];
        REAL shift_len = -2. * c->g;
        size_t i = c->cell / P->njk, rest = c->cell - i * P->njk, j = rest / (size_t)P->nk,
               k = rest - j * (size_t)P->nk;
        REAL xc = P->xccmin + (REAL)i * P->dx;
        REAL yc = P->yccmin + (REAL)j * P->dy;
        REAL zc = P->zccmin + (REAL)k * P->dz;
        REAL mx = xc + shift_len * gg[0];
        REAL my = yc + shift_len * gg[1];
        REAL mz = zc + shift_len * gg[2];
        wset_t ws;
        int r = cc_interp(P, mx, my, mz, &ws);
        if (r < 0) return -1;
        c->ex_ok = 0;
        /* weightedset.h:623-673 tranferToFixedArrays(8); the reference pads with an
         * out-of-range read (v[v.size()]), here the first index is used, weight 0 */
        if (r == 1 && ws.n > 0 && ws.n <= 8) {
          for (size_t a = 0; a < ws.n; ++a) {
            c->donor[a] = ws.idx[a];
            c->weight[a] = ws.w[a];
          }
          for (size_t a = ws.n; a < 8; ++a) {
            c->donor[a] = ws.idx[0];
            c->weight[a] = 0;
          }
          c->ex_ok = 1;
        }
      }
      size_t m = 0;
      for (size_t ll = 0; ll < L->n; ++ll)
        if (L->v[ll].ex_ok) L->v[m++] = L->v[ll];
      L->n = m;
    }
    for (int layer = 0; layer < n_outer; ++layer) {
      lsvec_t* L = &P->outer[layer];
      for (size_t ll = 0; ll < L->n; ++ll) {
        lscell_t* c = &L->v[ll];
        REAL gg[3];
        nabla(P, var, c->cell, gg);
        REAL len = 0;
        for (int a = 0; a < 3; ++a) len += gg[a] * gg[a];
        len = RSQRT(len);
        for (int a = 0; a < 3; ++a) gg[a] /= len;
        c->gx = gg[0];
        c->gy = gg[1];
        c->gz = gg[2];
      }
    }
  }
  return 0;
}

const REAL* F(grid_levelset_g)(const GRID* g, int p) { return g->p[p].G; }

int F(grid_ls_num_cells)(const GRID* g, int p, int layer, int outer) {
  const opatch* P = &g->p[p];
  if (outer) return layer < P->n_outer ? (int)P->outer[layer].n : 0;
  return layer < P->n_inner ? (int)P->inner[layer].n : 0;
}

void F(grid_ls_cells)(const GRID* g, int p, int layer, int outer, int64_t* cells, int64_t* donors, REAL* gvals,
                      REAL* weights) {
  const opatch* P = &g->p[p];
  const lsvec_t* L = outer ? &P->outer[layer] : &P->inner[layer];
  for (size_t a = 0; a < L->n; ++a) {
    cells[a] = (int64_t)L->v[a].cell;
    gvals[4 * a + 0] = L->v[a].g;
    gvals[4 * a + 1] = L->v[a].gx;
    gvals[4 * a + 2] = L->v[a].gy;
    gvals[4 * a + 3] = L->v[a].gz;
    for (int s = 0; s < 8; ++s) {
      donors[8 * a + s] = (int64_t)L->v[a].donor[s];
      weights[8 * a + s] = L->v[a].weight[s];
    }
  }
}

void F(grid_set_levelset_op)(GRID* g, int op) { g->ls_op = op == 1 ? 1 : 0; }

/* patch.cpp:733-746: squares, their maximum and their sum in REAL, over every
 * cell of the patch, then the square roots in REAL */
void F(grid_variable_difference)(const GRID* g, int p, int f1, int v1, int f2, int v2, double* max_norm,
                                 double* l2_norm) {
  const opatch* P = &g->p[p];
  const REAL* a = fieldp((opatch*)P, f1) + (size_t)v1 * P->N;
  const REAL* b = fieldp((opatch*)P, f2) + (size_t)v2 * P->N;
  REAL mx = 0., l2 = 0.;
  for (size_t i = 0; i < P->N; ++i) {
    REAL d = b[i] - a[i];
    REAL diff = d * d;
    mx = mx > diff ? mx : diff;
    l2 += diff;
  }
  *max_norm = (double)RSQRT(mx);
  *l2_norm = (double)RSQRT(l2);
}

/* cpu_levelsetobjectbc.h:58-90 + lsobccompressibleeulerop.h:49-144 (inner cells;
 * operateOuter is a no-op), or with ls_op 1 lsobccompressibleswallop.h:48-116,
 * whose access is the same mirror sum.  The flat order is patch, layer, list
 * position (levelsetobjectbc.cpp:38-134). */
void F(grid_levelset_fill)(GRID* g, int field, int abuse_field) {
  const REAL relax = 0.5;
  for (int p = 0; p < g->np; ++p) {
    opatch* P = &g->p[p];
    const REAL* f = fieldp(P, field);
    REAL* acc = fieldp(P, abuse_field);
    const size_t N = P->N;
    for (int layer = 0; layer < P->n_inner; ++layer) {
      const lsvec_t* L = &P->inner[layer];
      for (size_t ll = 0; ll < L->n; ++ll) {
        size_t l = L->v[ll].cell;
        for (int v = 0; v < 5; ++v) acc[v * N + l] = 0.;
        for (int s = 0; s < 8; ++s) {
          size_t dn = L->v[ll].donor[s];
          REAL w = L->v[ll].weight[s];
          for (int v = 0; v < 5; ++v) acc[v * N + l] += w * f[v * N + dn];
        }
      }
    }
  }
  for (int p = 0; p < g->np; ++p) {
    opatch* P = &g->p[p];
    REAL* f = fieldp(P, field);
    const REAL* acc = fieldp(P, abuse_field);
    const size_t N = P->N;
    for (int layer = 0; layer < P->n_inner; ++layer) {
      const lsvec_t* L = &P->inner[layer];
      for (size_t ll = 0; ll < L->n; ++ll) {
        size_t l = L->v[ll].cell;
        if (g->ls_op == 1) {
          /* lsobccompressibleswallop.h:82-116: density and total energy from the
           * mirror sum, momentum reversed and relaxed; the double literal makes
           * (1. - relax) * rhou a double sum, stored in REAL */
          f[0 * N + l] = acc[0 * N + l];
          for (int a = 1; a <= 3; ++a) f[a * N + l] = relax * (-acc[a * N + l]) + (1. - relax) * f[a * N + l];
          f[4 * N + l] = acc[4 * N + l];
          continue;
        }
        REAL gxn = L->v[ll].gx, gyn = L->v[ll].gy, gzn = L->v[ll].gz;
        REAL rho_uvw_n = acc[1 * N + l] * gxn + acc[2 * N + l] * gyn + acc[3 * N + l] * gzn;
        REAL eps = 1.e-6;
        REAL hard[3], cur[3], damp[3];
        REAL gn[3] = {gxn, gyn, gzn};
        for (int a = 0; a < 3; ++a) {
          hard[a] = acc[(1 + a) * N + l] - 2. * rho_uvw_n * gn[a];
          cur[a] = f[(1 + a) * N + l];
          REAL diff = cur[a] - hard[a];
          damp[a] = RFABS(diff) / (RFABS(cur[a]) + eps);
          if (damp[a] > (1. - relax)) damp[a] = 1. - relax;
        }
        for (int a = 0; a < 3; ++a) f[(1 + a) * N + l] = (1. - damp[a]) * hard[a] + damp[a] * cur[a];
        f[0 * N + l] = acc[0 * N + l];
        f[4 * N + l] = acc[4 * N + l];
      }
    }
  }
}

/* rungekutta.cpp:27-42 with TimeIntegration (timeintegration.cpp:28-55); the
 * level-set fill runs as the post operation when a level set is set up. */
void F(grid_rk_step)(GRID* g, REAL dt, int n_alpha, const REAL* alpha) {
  F(grid_exchange)(g, 0);
  F(grid_copy_field)(g, 0, 1);
  for (int s = 0; s < n_alpha; ++s) {
    if (s > 0) F(grid_exchange)(g, 0);
    F(grid_compute)(g, alpha[s] * dt);
    F(grid_exchange)(g, 0);
    if (g->have_ls) F(grid_levelset_fill)(g, 0, 2);
  }
}

/* FP64 sums of the conserved variables times cell volume over the core cells */
void F(grid_core_sums)(const GRID* g, int field, double* out) {
  for (int v = 0; v < 5; ++v) out[v] = 0;
  for (int p = 0; p < g->np; ++p) {
    const opatch* P = &g->p[p];
    const REAL* f = fieldp(P, field);
    double dV = (double)P->dx * (double)P->dy * (double)P->dz;
    for (int v = 0; v < 5; ++v) {
      double s = 0;
      for (size_t i = (size_t)P->seek[0]; i < (size_t)P->ni - P->seek[1]; ++i)
        for (size_t j = (size_t)P->seek[2]; j < (size_t)P->nj - P->seek[3]; ++j)
          for (size_t k = (size_t)P->seek[4]; k < (size_t)P->nk - P->seek[5]; ++k) s += (double)f[v * P->N + IDX(P, i, j, k)];
      out[v] += s * dV;
    }
  }
}
