// drnhip.hip -- libdrnhip.so: context, device memory, work lists, launches and
// the C ABI declared in include/drnhip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/drnhip.h"
#include "drn_kernels.h"

using namespace drn;

namespace {

constexpr int NBUF = 3;
constexpr int ALIGN_CELLS = 64;

struct ExchangeTables {
  // reference form, per receiver patch (local cell indices)
  std::vector<int32_t> donor_patch;
  std::vector<int64_t> donor_info;  // 4 per donor: nrec, stride, rec_start, wi_start
  std::vector<int64_t> rec_unique, rec_concat, donor_index;
  std::vector<double> donor_weight;
  // 9 per donor (row-major), the turn of the momentum from the donor's frame
  // into the receiver's (donor_t axx..azz); empty when every donor is aligned
  std::vector<double> turn;
};

// a turn matrix that changes no finite vector: exact 1 on the diagonal, zeros off it
static bool is_identity_turn(const double* a) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      if (a[3 * r + c] != (r == c ? 1.0 : 0.0)) return false;
  return true;
}

struct LsTables {
  std::vector<int64_t> cells, donors;
  std::vector<double> normals, weights;
};

struct TimerSlot {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> used, pool;
};

}  // namespace

struct drn_ctx {
  int device = 0;
  int n_cu = 0;  // compute units of the device (persistent launch size)
  int prec = DRN_F64;
  hipStream_t stream = nullptr;
  double* scratch = nullptr;  // device address of the file-scope `scratch` array on this device
  std::string err;
  int recon = DRN_RECON_UPWIND2, lim = DRN_LIM_VANALBADA;
  bool visc = false;  // CompressibleViscFlux added to the interior face fluxes
  // per-cell flags (DRN_CELL_*), one byte per stored cell in the state layout;
  // null until a patch gets flags, flags_on while any cell has one
  uint8_t* d_flags = nullptr;
  bool flags_on = false;
  std::vector<uint8_t> flags_any;  // per patch: whether it has a nonzero flag

  std::vector<drn_patch_desc> descs;
  std::vector<uint64_t> offs;
  uint64_t vstride = 0;  // cells per variable block (all patches, padded)
  bool laid_out = false;
  void* buf[NBUF] = {nullptr, nullptr, nullptr};
  int map[NBUF] = {0, 1, 2};  // logical field -> physical buffer

  PatchDev* d_pd = nullptr;
  // the descriptors of the RK step: faces whose boundary cells are all exchange
  // receivers carry BC_DEAD (their functor's contribution is overwritten by the
  // exchange after the stage); null when no face is dead
  PatchDev* d_pd_rk = nullptr;
  int n_dead_faces = 0;
  const PatchDev* pd_stage = nullptr;  // the descriptors of the stage being launched
  std::vector<TileCell> tiles;
  std::vector<int> tile_first;  // per patch, size np+1
  TileCell* d_tiles = nullptr;
  TileCell* d_tiles_xcd = nullptr;     // the list in 8 contiguous segments, one per XCD
  TileCell* d_tiles_xcd_pp = nullptr;  // the same per patch (same tile_first offsets)
  // 252-cell items of the z-shared cell kernel (4 wavefronts x 63 cells)
  std::vector<TileCell> tz;
  std::vector<int> tz_first;
  TileCell* d_tz = nullptr;
  TileCell* d_tz_xcd = nullptr;
  TileCell* d_tz_chain = nullptr;  // z-shared list interleaved along the fused push chains (commit)
  TileCell* d_tz_xcd_pp = nullptr;
  // 2-plane x 252-cell items of the two-plane cell kernel
  // items of the x-shared cell kernels, XP planes x (4/XP) x 63 cells; xs[0]: XP = 2, xs[1]: XP = 4
  struct XsList {
    std::vector<TileCell> nat;
    std::vector<int> first;
    TileCell *d = nullptr, *d_xcd = nullptr, *d_xcd_pp = nullptr, *d_chain = nullptr;
  } xs[2];
  std::vector<TileCell> tx2;
  std::vector<int> tx2_first;
  TileCell* d_tx2 = nullptr;
  TileCell* d_tx2_xcd = nullptr;
  TileCell* d_tx2_xcd_pp = nullptr;
  // 504-cell items of the y/z-shared cell kernel (8 wavefronts x 63 cells)
  std::vector<TileCell> tyz;
  std::vector<int> tyz_first;
  TileCell* d_tyz = nullptr;
  TileCell* d_tyz_xcd = nullptr;
  TileCell* d_tyz_xcd_pp = nullptr;
  double* d_part = nullptr;  // reduction partials (5 per tile)
  double* d_out = nullptr;
  void* d_post = nullptr;    // post-processing output of one patch
  size_t post_bytes = 0;
  int tiling = 11;     // batched form
  int tiling_pp = 11;  // per-patch form
  int arith = -1;  // DRN_ARITH_*; -1 = precision default
  // marching work-lists: batched (launch order) and per-patch (natural order, patch-contiguous)
  std::vector<TileMarch> mt_pp;
  std::vector<int> mt_first;  // per patch, size np+1 (into mt_pp)
  TileMarch* d_mt_pp = nullptr;
  TileMarch* d_mt_launch = nullptr;
  int n_mt = 0;
  int lds_cells = 0, lds_faces = 0, lds_sy = 0, lds_sz = 0;
  std::vector<TileWave> wt_pp;  // wave-tile items per patch (per-patch form)
  std::vector<int> wt_first;    // per patch, size np+1 (into wt_pp)
  TileWave* d_wt_pp = nullptr;
  TileWave* d_wt_launch = nullptr;  // batched form, in launch order
  int n_wt = 0;
  std::vector<int> thin;      // patches routed to the cell kernel
  std::vector<TileCell> thin_tiles;
  TileCell* d_thin_tiles = nullptr;

  // exchange: the full receiver lists, and the lists without the z-face
  // receivers that the stage kernel fills itself when the exchange is fused
  std::map<int, ExchangeTables> xt;
  struct ExchangeLists {
    struct SingleClass {
      int stride;
      uint32_t n;
      uint32_t *rec, *idx;
      void* w;
    };
    std::vector<SingleClass> single;
    // several donor patches per receiver (k_exchange_multi): receivers sorted by
    // entry count, entries interleaved per wavefront of 64 receivers
    uint32_t multi_n = 0;
    uint32_t *m_rec = nullptr, *m_cnt = nullptr, *m_tbase = nullptr, *m_ent = nullptr;
    void* m_w = nullptr;
    uint32_t* m_et = nullptr;  // per entry closing a donor sum: 0 = aligned, t > 0 = turn matrix t - 1 of m_turn
    void* m_turn = nullptr;    // 9 values per turn matrix, in the state precision
  };
  ExchangeLists xfull, xpart;
  bool have_exchange = false;
  std::vector<PatchDev> pd_host;  // the device patch descriptors as uploaded
  bool fuse = true;               // drn_set_exchange_fusion
  bool push_now = false;          // the stage being launched fills the fused receivers
  int64_t n_fused = 0;            // receivers the stage kernel fills when fused

  // level set
  std::map<int, LsTables> lt;
  uint32_t ls_n = 0;
  int ls_op = DRN_LS_SLIP;  // the operator of the fill
  uint32_t *ls_cells = nullptr, *ls_donors = nullptr;
  void *ls_w = nullptr, *ls_nrm = nullptr, *ls_acc = nullptr;

  // timing
  bool timing = false;
  TimerSlot timers[4];

  // Captured step cycles, one per starting field map: a graph holds the
  // physical buffers of the map it was captured from, so it is replayed only
  // from that map.
  struct GraphEntry {
    int map[NBUF];
    double dt;
    std::vector<double> alpha;
    int steps;      // steps per replay (the cycle of the field-map permutation)
    int exch_end;   // exch_buf after one replay
    hipGraphExec_t exec;
  };
  std::vector<GraphEntry> graphs;
  // The physical buffer whose receivers hold the exchange of its current
  // donor values (-1: none).  Every write to a buffer clears it for that
  // buffer; a step without post operation skips its leading exchange
  // (rungekutta.cpp:30) only when it is set for the state buffer.
  int exch_buf = -1;

  size_t esize() const { return prec == DRN_F32 ? 4 : 8; }
};

#define HIPCHK(ctx, expr)                                                            \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) {                                                          \
      (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                \
      return DRN_E_HIP;                                                              \
    }                                                                                \
  } while (0)

static void drop_graph(drn_ctx* ctx) {
  for (auto& g : ctx->graphs) (void)hipGraphExecDestroy(g.exec);
  ctx->graphs.clear();
}

// buffer b is written: its receivers no longer hold the exchange
static void wrote(drn_ctx* ctx, int b) {
  if (ctx->exch_buf == b) ctx->exch_buf = -1;
}

static int fail(drn_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

template <typename T>
static hipError_t dalloc(T** p, size_t n) {
  return hipMalloc((void**)p, n * sizeof(T) + 16);
}

static void dfree(void* p) {
  if (p) (void)hipFree(p);
}

// ---------------------------------------------------------------------------
// timing helpers
// ---------------------------------------------------------------------------

static void timer_begin(drn_ctx* ctx, int which, hipEvent_t* stop_out) {
  *stop_out = nullptr;
  if (!ctx->timing) return;
  TimerSlot& s = ctx->timers[which];
  std::pair<hipEvent_t, hipEvent_t> ev;
  if (!s.pool.empty()) {
    ev = s.pool.back();
    s.pool.pop_back();
  } else {
    (void)hipEventCreate(&ev.first);
    (void)hipEventCreate(&ev.second);
  }
  (void)hipEventRecord(ev.first, ctx->stream);
  s.used.push_back(ev);
  *stop_out = ev.second;
}

static void timer_end(drn_ctx* ctx, hipEvent_t stop) {
  if (stop) (void)hipEventRecord(stop, ctx->stream);
}

// ---------------------------------------------------------------------------
// launches
// ---------------------------------------------------------------------------

constexpr int MARCH_BS = 256;

// tiling choices: kernel, i-range length (0 = whole patch) and work-list order
enum {
  K_MARCH = 0,
  K_CELL = 1,
  K_COLUMN = 2,
  K_CELL_SLOPE = 3,
  K_WTILE = 4,
  K_CELLZ = 5,
  K_CELLYZ = 6,
  K_CELLX2 = 7,
  K_CELLZP = 8,  // K_CELLZ items, one persistent launch
  K_XM = 9,      // the z-shared cell kernel marching its (j,k) cells over an i-range
  K_XM8 = 10,    // K_XM capped at 8 waves per SIMD (FP32 fast arithmetic; else K_XM)
  K_XM7 = 11,    // K_XM capped at 7 waves per SIMD (FP32 fast arithmetic; else K_XM)
  K_XML = 12,    // K_XM with the x slope and x face flux carried in LDS
  K_CELLXS2 = 13,  // the z-shared cell kernel over 2 planes per item, x faces shared through LDS
  K_CELLXS4 = 14   // the same over 4 planes per item
};
struct TilingChoice {
  int irange;  // planes per work item, 0 = whole patch
  int xcd;     // 1 = segment the work list so that neighbouring chunks share an XCD;
               // 2 = the same over patch planes interleaved along the chains of
               //     fused z-face pushes (z-shared cell kernel, batched launch)
  int kind;    // K_MARCH (LDS planes, or the slope form in fast mode), K_CELL, K_COLUMN
};
static const TilingChoice kTilings[] = {
    {0, 0, K_MARCH},  {0, 1, K_MARCH},  {32, 0, K_MARCH},  {32, 1, K_MARCH}, {16, 0, K_MARCH},
    {16, 1, K_MARCH}, {8, 0, K_MARCH},  {8, 1, K_MARCH},   {4, 0, K_MARCH},  {4, 1, K_MARCH},
    {2, 0, K_MARCH},  {0, 0, K_CELL},   {0, 0, K_COLUMN},  {0, 1, K_COLUMN}, {32, 1, K_COLUMN},
    {16, 0, K_COLUMN}, {16, 1, K_COLUMN}, {8, 1, K_COLUMN}, {4, 1, K_COLUMN}, {0, 0, K_CELL_SLOPE},
    {0, 1, K_CELL},    {0, 1, K_CELL_SLOPE}, {0, 1, K_WTILE}, {16, 1, K_WTILE}, {8, 1, K_WTILE},
    {4, 1, K_WTILE},   {8, 0, K_WTILE},      {0, 1, K_CELLZ},  {0, 0, K_CELLZ},  {0, 1, K_CELLYZ},
    {0, 0, K_CELLYZ},  {0, 1, K_CELLX2},     {0, 0, K_CELLX2}, {0, 1, K_CELLZP}, {0, 0, K_CELLZP},
    {0, 2, K_CELLZ},   {0, 1, K_XM},         {32, 1, K_XM},    {16, 1, K_XM},    {8, 1, K_XM},
    {0, 1, K_XM8},     {16, 1, K_XM8},       {0, 1, K_XM7},    {16, 1, K_XM7},   {0, 1, K_XML},
    {16, 1, K_XML},    {8, 1, K_XML},        {0, 2, K_CELLXS4}, {0, 1, K_CELLXS4}, {0, 2, K_CELLXS2},
    {0, 1, K_CELLXS2}};
static const int kNumTilings = sizeof(kTilings) / sizeof(kTilings[0]);

// the column kernel covers the fast MUSCL forms with a symmetric limited slope
static bool column_capable(const drn_ctx* ctx);

template <typename T, int RECON, int LIM, bool FAST>
static void launch_march_t(drn_ctx* ctx, const TileMarch* tiles, int ntiles, const T* src, const T* base, T* dst,
                           T factor, bool column) {
  if (ntiles <= 0) return;
  if constexpr (FAST && RECON == RECON_UPWIND2 && (LIM == LIM_VANALBADA || LIM == LIM_MINMOD)) {
    if (column) {
      hipLaunchKernelGGL((k_stage_column<T, LIM>), dim3(ntiles), dim3(256), 0, ctx->stream, ctx->pd_stage, tiles, src,
                         base, dst, ctx->vstride, factor);
      return;
    }
    // slope form: per-cell limited slopes shared by both faces of the cell
    const size_t lds = (size_t)(10 * ctx->lds_cells + 5 * (ctx->lds_sy + ctx->lds_sz + ctx->lds_faces)) * sizeof(T);
    hipLaunchKernelGGL((k_stage_slope<T, LIM, MARCH_BS>), dim3(ntiles), dim3(MARCH_BS), lds, ctx->stream, ctx->pd_stage,
                       tiles, src, base, dst, ctx->vstride, factor, ctx->lds_cells, ctx->lds_faces, ctx->lds_sy,
                       ctx->lds_sz);
  } else {
    const size_t lds = (size_t)5 * (ctx->lds_cells + ctx->lds_faces) * sizeof(T);
    hipLaunchKernelGGL((k_stage_march<T, RECON, LIM, FAST, MARCH_BS>), dim3(ntiles), dim3(MARCH_BS), lds, ctx->stream,
                       ctx->pd_stage, tiles, src, base, dst, ctx->vstride, factor, ctx->lds_cells, ctx->lds_faces);
  }
}

template <typename T, int RECON, int LIM, bool FAST>
static void launch_cell_t(drn_ctx* ctx, const TileCell* tiles, int ntiles, const T* src, const T* base, T* dst,
                          T factor, bool slope) {
  if (ntiles <= 0) return;
  if constexpr (FAST && RECON == RECON_UPWIND2 && (LIM == LIM_VANALBADA || LIM == LIM_MINMOD)) {
    if (slope) {
      hipLaunchKernelGGL((k_stage_cell_slope<T, LIM, 256>), dim3(ntiles), dim3(256), 0, ctx->stream, ctx->pd_stage, tiles,
                         src, base, dst, ctx->vstride, factor);
      return;
    }
  }
  {
    hipLaunchKernelGGL((k_stage_cell<T, RECON, LIM, FAST, 256>), dim3(ntiles), dim3(256), 0, ctx->stream, ctx->pd_stage,
                       tiles, src, base, dst, ctx->vstride, factor, (int)ctx->push_now);
  }
}

template <typename T, int RECON, int LIM, bool FAST>
static void launch_wtile_t(drn_ctx* ctx, const TileWave* items, int n, const T* src, const T* base, T* dst, T factor) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_stage_wtile<T, RECON, LIM, FAST>), dim3((n + 3) / 4), dim3(256), 0, ctx->stream, ctx->pd_stage,
                     items, n, src, base, dst, ctx->vstride, factor);
}

// one launch of either kernel over a tile range (march tiles or cell tiles)
// kind: K_CELL runs k_stage_cell on TileCell items, otherwise the marching
// kernels (column kernel when kind == K_COLUMN and the operator allows it)
template <typename T, int RECON, int LIM, bool FAST>
static void launch_range(drn_ctx* ctx, int kind, const void* tiles, int ntiles, const T* src, const T* base, T* dst,
                         T factor) {
  if (kind == K_CELL || kind == K_CELL_SLOPE)
    launch_cell_t<T, RECON, LIM, FAST>(ctx, (const TileCell*)tiles, ntiles, src, base, dst, factor,
                                       kind == K_CELL_SLOPE);
  else if (kind == K_WTILE)
    launch_wtile_t<T, RECON, LIM, FAST>(ctx, (const TileWave*)tiles, ntiles, src, base, dst, factor);
  else if (kind == K_CELLXS2 || kind == K_CELLXS4) {
    if (ntiles > 0) {
      if (kind == K_CELLXS2)
        hipLaunchKernelGGL((k_stage_cell_xs<T, RECON, LIM, FAST, 2>), dim3(ntiles), dim3(256), 0, ctx->stream,
                           ctx->pd_stage, (const TileCell*)tiles, src, base, dst, ctx->vstride, factor,
                           (int)ctx->push_now);
      else
        hipLaunchKernelGGL((k_stage_cell_xs<T, RECON, LIM, FAST, 4>), dim3(ntiles), dim3(256), 0, ctx->stream,
                           ctx->pd_stage, (const TileCell*)tiles, src, base, dst, ctx->vstride, factor,
                           (int)ctx->push_now);
    }
  } else if (kind == K_CELLX2) {
    if (ntiles > 0)
      hipLaunchKernelGGL((k_stage_cell_x2<T, RECON, LIM, FAST>), dim3(ntiles), dim3(256), 0, ctx->stream, ctx->pd_stage,
                         (const TileCell*)tiles, src, base, dst, ctx->vstride, factor);
  } else if (kind == K_CELLYZ) {
    if (ntiles > 0)
      hipLaunchKernelGGL((k_stage_cell_yz<T, RECON, LIM, FAST>), dim3(ntiles), dim3(512), 0, ctx->stream, ctx->pd_stage,
                         (const TileCell*)tiles, src, base, dst, ctx->vstride, factor);
  } else if (kind == K_CELLZ && ctx->flags_on) {
    if (ntiles > 0) {
      if (ctx->visc)
        hipLaunchKernelGGL((k_stage_cell_flags<T, RECON, LIM, FAST, true>), dim3(ntiles), dim3(256), 0, ctx->stream,
                           ctx->pd_stage, (const TileCell*)tiles, src, base, dst, ctx->vstride, factor,
                           (int)ctx->push_now, (const uint8_t*)ctx->d_flags);
      else
        hipLaunchKernelGGL((k_stage_cell_flags<T, RECON, LIM, FAST, false>), dim3(ntiles), dim3(256), 0, ctx->stream,
                           ctx->pd_stage, (const TileCell*)tiles, src, base, dst, ctx->vstride, factor,
                           (int)ctx->push_now, (const uint8_t*)ctx->d_flags);
    }
  } else if (kind == K_CELLZ) {
    if (ntiles > 0 && ctx->visc) {
      if constexpr (sizeof(T) == 4)
        hipLaunchKernelGGL((k_stage_cell_visc_f32<RECON, LIM, FAST>), dim3(ntiles), dim3(256), 0, ctx->stream,
                           ctx->pd_stage, (const TileCell*)tiles, src, base, dst, ctx->vstride, factor, (int)ctx->push_now);
      else
        hipLaunchKernelGGL((k_stage_cell<T, RECON, LIM, FAST, 256, true, true>), dim3(ntiles), dim3(256), 0,
                           ctx->stream, ctx->pd_stage, (const TileCell*)tiles, src, base, dst, ctx->vstride, factor,
                           (int)ctx->push_now);
    } else if (ntiles > 0)
      hipLaunchKernelGGL((k_stage_cell<T, RECON, LIM, FAST, 256, true>), dim3(ntiles), dim3(256), 0, ctx->stream,
                         ctx->pd_stage, (const TileCell*)tiles, src, base, dst, ctx->vstride, factor, (int)ctx->push_now);
  } else if (kind == K_XM || kind == K_XM8 || kind == K_XM7 || kind == K_XML) {
    if (ntiles > 0) {
      auto kern = k_stage_xm<T, RECON, LIM, FAST>;
      if (kind == K_XML) kern = k_stage_xml<T, RECON, LIM, FAST>;
      if constexpr (FAST && sizeof(T) == 4) {
        if (kind == K_XM8) kern = k_stage_xm8<T, RECON, LIM, FAST>;
        if (kind == K_XM7) kern = k_stage_xm7<T, RECON, LIM, FAST>;
      }
      hipLaunchKernelGGL(kern, dim3(ntiles), dim3(256), 0, ctx->stream, ctx->pd_stage, (const TileMarch*)tiles, src, base,
                         dst, ctx->vstride, factor, (int)ctx->push_now);
    }
  } else if (kind == K_CELLZP) {
    if (ntiles > 0) {
      // resident workgroups per CU of this instantiation (a property of the
      // kernel and the target), queried once; the initialisation of a local
      // static is thread-safe, so contexts driven from several threads agree
      static const int per_cu = [] {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_stage_cell_persist<T, RECON, LIM, FAST>, 256, 0) !=
                hipSuccess ||
            n <= 0)
          n = 1;
        return n;
      }();
      // a multiple of 8 keeps every workgroup on the XCD of its first item
      int g = std::max(8, (ctx->n_cu * per_cu) / 8 * 8);
      if (g > ntiles) g = ntiles;
      hipLaunchKernelGGL((k_stage_cell_persist<T, RECON, LIM, FAST>), dim3(g), dim3(256), 0, ctx->stream, ctx->pd_stage,
                         (const TileCell*)tiles, ntiles, src, base, dst, ctx->vstride, factor, (int)ctx->push_now);
    }
  }
  else
    launch_march_t<T, RECON, LIM, FAST>(ctx, (const TileMarch*)tiles, ntiles, src, base, dst, factor,
                                        kind == K_COLUMN);
}

template <typename T, bool FAST>
static void launch_dispatch(drn_ctx* ctx, int kind, const void* tiles, int ntiles, const T* src, const T* base, T* dst,
                            T factor) {
  if (ctx->recon == DRN_RECON_UPWIND1)
    launch_range<T, RECON_UPWIND1, LIM_SECONDORDER, FAST>(ctx, kind, tiles, ntiles, src, base, dst, factor);
  else if (ctx->lim == DRN_LIM_VANALBADA)
    launch_range<T, RECON_UPWIND2, LIM_VANALBADA, FAST>(ctx, kind, tiles, ntiles, src, base, dst, factor);
  else if (ctx->lim == DRN_LIM_MINMOD)
    launch_range<T, RECON_UPWIND2, LIM_MINMOD, FAST>(ctx, kind, tiles, ntiles, src, base, dst, factor);
  else
    launch_range<T, RECON_UPWIND2, LIM_SECONDORDER, FAST>(ctx, kind, tiles, ntiles, src, base, dst, factor);
}

static bool column_capable(const drn_ctx* ctx) {
  const bool fast = ctx->arith < 0 ? ctx->prec == DRN_F32 : ctx->arith == DRN_ARITH_FAST;
  return fast && ctx->recon == DRN_RECON_UPWIND2 && (ctx->lim == DRN_LIM_VANALBADA || ctx->lim == DRN_LIM_MINMOD);
}

// effective kernel kind of a tiling choice in the current arithmetic/operator;
// with the viscous flux every choice runs the z-shared cell kernel on its list
// order (the other kernels carry the Euler flux only)
static int effective_kind(const drn_ctx* ctx, int choice) {
  if (ctx->visc || ctx->flags_on) return K_CELLZ;
  const int k = kTilings[choice].kind;
  if (k == K_COLUMN && !column_capable(ctx)) return K_MARCH;
  if (k == K_CELL_SLOPE && !column_capable(ctx)) return K_CELL;
  const bool fast32 = ctx->prec == DRN_F32 && (ctx->arith < 0 || ctx->arith == DRN_ARITH_FAST);
  if ((k == K_XM8 || k == K_XM7) && !fast32) return K_XM;
  return k;
}

static bool fast_math(const drn_ctx* ctx) {
  return ctx->arith < 0 ? ctx->prec == DRN_F32 : ctx->arith == DRN_ARITH_FAST;
}

template <typename T>
static void stage_t(drn_ctx* ctx, int psrc, int pbase, int pdst, T factor, int launch) {
  const T* src = (const T*)ctx->buf[psrc];
  const T* base = (const T*)ctx->buf[pbase];
  T* dst = (T*)ctx->buf[pdst];
  const bool fast = fast_math(ctx);
  auto go = [&](int kind, const void* tl, int n) {
    if (fast)
      launch_dispatch<T, true>(ctx, kind, tl, n, src, base, dst, factor);
    else
      launch_dispatch<T, false>(ctx, kind, tl, n, src, base, dst, factor);
  };
  if (launch == DRN_LAUNCH_BATCHED) {
    const int kind = effective_kind(ctx, ctx->tiling);
    if (kind == K_CELLXS2 || kind == K_CELLXS4) {
      const auto& L = ctx->xs[kind == K_CELLXS4 ? 1 : 0];
      const int xo = kTilings[ctx->tiling].xcd;
      go(kind, xo == 2 && L.d_chain ? L.d_chain : (xo ? L.d_xcd : L.d), (int)L.nat.size());
    } else if (kind == K_CELLX2) {
      go(K_CELLX2, kTilings[ctx->tiling].xcd ? ctx->d_tx2_xcd : ctx->d_tx2, (int)ctx->tx2.size());
    } else if (kind == K_CELLYZ) {
      go(K_CELLYZ, kTilings[ctx->tiling].xcd ? ctx->d_tyz_xcd : ctx->d_tyz, (int)ctx->tyz.size());
    } else if (kind == K_CELLZ || kind == K_CELLZP) {
      const int xo = kTilings[ctx->tiling].xcd;
      const TileCell* tl = xo == 2 && ctx->d_tz_chain ? ctx->d_tz_chain : (xo ? ctx->d_tz_xcd : ctx->d_tz);
      go(kind, tl, (int)ctx->tz.size());
    } else if (kind == K_CELL || kind == K_CELL_SLOPE) {
      go(kind, kTilings[ctx->tiling].xcd ? ctx->d_tiles_xcd : ctx->d_tiles, (int)ctx->tiles.size());
    } else if (kind == K_WTILE) {
      go(K_WTILE, ctx->d_wt_launch, ctx->n_wt);
      go(K_CELL, ctx->d_thin_tiles, (int)ctx->thin_tiles.size());
    } else {
      go(kind, ctx->d_mt_launch, ctx->n_mt);
      go(K_CELL, ctx->d_thin_tiles, (int)ctx->thin_tiles.size());
    }
  } else {
    const int kind = effective_kind(ctx, ctx->tiling_pp);
    const int np = (int)ctx->descs.size();
    for (int p = 0; p < np; ++p) {
      const int m0 = ctx->mt_first[p], mn = ctx->mt_first[p + 1] - m0;
      const int w0 = ctx->wt_first[p], wn = ctx->wt_first[p + 1] - w0;
      if (kind == K_CELLXS2 || kind == K_CELLXS4) {
        const auto& L = ctx->xs[kind == K_CELLXS4 ? 1 : 0];
        const int z0 = L.first[p];
        go(kind, (kTilings[ctx->tiling_pp].xcd ? L.d_xcd_pp : L.d) + z0, L.first[p + 1] - z0);
      } else if (kind == K_CELLX2) {
        const int z0 = ctx->tx2_first[p];
        const TileCell* tl = (kTilings[ctx->tiling_pp].xcd ? ctx->d_tx2_xcd_pp : ctx->d_tx2) + z0;
        go(K_CELLX2, tl, ctx->tx2_first[p + 1] - z0);
      } else if (kind == K_CELLYZ) {
        const int z0 = ctx->tyz_first[p];
        const TileCell* tl = (kTilings[ctx->tiling_pp].xcd ? ctx->d_tyz_xcd_pp : ctx->d_tyz) + z0;
        go(K_CELLYZ, tl, ctx->tyz_first[p + 1] - z0);
      } else if (kind == K_CELLZ || kind == K_CELLZP) {
        const int z0 = ctx->tz_first[p];
        const TileCell* tl = (kTilings[ctx->tiling_pp].xcd ? ctx->d_tz_xcd_pp : ctx->d_tz) + z0;
        go(kind, tl, ctx->tz_first[p + 1] - z0);
      } else if (kind == K_WTILE && wn > 0) {
        go(K_WTILE, ctx->d_wt_pp + w0, wn);
      } else if (kind == K_CELL || kind == K_CELL_SLOPE || kind == K_WTILE || mn == 0) {
        const int t0 = ctx->tile_first[p];
        const TileCell* tl = (kTilings[ctx->tiling_pp].xcd ? ctx->d_tiles_xcd_pp : ctx->d_tiles) + t0;
        go(kind == K_CELL_SLOPE ? K_CELL_SLOPE : K_CELL, tl, ctx->tile_first[p + 1] - t0);
      } else {
        go(kind, ctx->d_mt_pp + m0, mn);
      }
    }
  }
}

// The cell kernels (choices of kind K_CELL, K_CELLZ, K_CELLZP, K_XM*) can fill the
// fused z-face receivers themselves; the other kinds leave them to the exchange.
static bool push_kind(int kind) {
  return kind == K_CELL || kind == K_CELLZ || kind == K_CELLZP || kind == K_XM || kind == K_XM8 || kind == K_XM7 ||
         kind == K_XML || kind == K_CELLXS2 || kind == K_CELLXS4;
}

// whether a stage of this launch form fills the fused receivers (and the
// exchange after it then runs on the lists without them)
static bool push_active(const drn_ctx* ctx, int launch) {
  if (!ctx->fuse || !ctx->have_exchange || ctx->n_fused == 0) return false;
  return push_kind(effective_kind(ctx, launch == DRN_LAUNCH_BATCHED ? ctx->tiling : ctx->tiling_pp));
}

// rk: a stage of the RK step (or the autotuner's stand-in for one), whose
// exchange follows at once, so that dead faces may skip their functor
static int do_stage(drn_ctx* ctx, int psrc, int pbase, int pdst, double factor, int launch, bool push = false,
                    bool rk = false) {
  wrote(ctx, pdst);
  hipEvent_t stop;
  timer_begin(ctx, 0, &stop);
  ctx->push_now = push;
  ctx->pd_stage = rk && ctx->fuse && ctx->d_pd_rk ? ctx->d_pd_rk : ctx->d_pd;
  if (ctx->prec == DRN_F32)
    stage_t<float>(ctx, psrc, pbase, pdst, (float)factor, launch);
  else
    stage_t<double>(ctx, psrc, pbase, pdst, factor, launch);
  ctx->push_now = false;
  timer_end(ctx, stop);
  HIPCHK(ctx, hipGetLastError());
  return DRN_OK;
}

// marching work items of one patch for an i-range length L (0 = whole patch);
// i ranges outermost so that the chunks of one plane range are adjacent
static void march_items(const drn_patch_desc& d, int p, int L, int chunk, std::vector<TileMarch>& out) {
  const int njk = d.nj * d.nk;
  const int Lp = L > 0 ? L : d.ni;
  for (int i0 = 0; i0 < d.ni; i0 += Lp)
    for (int c0 = 0; c0 < njk; c0 += chunk) out.push_back(TileMarch{p, i0, std::min(d.ni, i0 + Lp), c0});
}

// wave-tile items of one patch for an i-range length L (0 = whole patch):
// 8x8-lane tiles advance 7x7 columns; k tiles innermost
static void wtile_items(const drn_patch_desc& d, int p, int L, std::vector<TileWave>& out) {
  const int Lp = L > 0 ? L : d.ni;
  for (int i0 = 0; i0 < d.ni; i0 += Lp)
    for (int j0 = 0; j0 < d.nj; j0 += 7)
      for (int k0 = 0; k0 < d.nk; k0 += 7)
        out.push_back(TileWave{p, (int16_t)i0, (int16_t)std::min(d.ni, i0 + Lp), (int16_t)j0, (int16_t)k0, 0});
}

// cells of the (j,k) plane per marching work item: the column kernel's four
// wavefronts own 63 cells each
static int chunk_of(const drn_ctx* ctx, int choice) {
  const int k = effective_kind(ctx, choice);
  return k == K_COLUMN || k == K_XM || k == K_XM8 || k == K_XM7 || k == K_XML ? 252 : MARCH_BS;
}

// build the marching work-lists of the current tiling choices of both launch forms
static int build_march(drn_ctx* ctx) {
  const TilingChoice tb = kTilings[ctx->tiling], tp = kTilings[ctx->tiling_pp];
  const int np = (int)ctx->descs.size();
  std::vector<TileMarch> nat;
  std::vector<TileWave> wnat;
  ctx->mt_pp.clear();
  ctx->mt_first.assign(np + 1, 0);
  ctx->wt_pp.clear();
  ctx->wt_first.assign(np + 1, 0);
  ctx->thin.clear();
  ctx->thin_tiles.clear();
  int lc = 0, lf = 0, lsy = 0, lsz = 0;
  for (int p = 0; p < np; ++p) {
    const auto& d = ctx->descs[p];
    ctx->mt_first[p] = (int)ctx->mt_pp.size();
    const int njk = d.nj * d.nk;
    int plc = 0, plf = 0, psy = 0, psz = 0;
    for (int c0 = 0; c0 < njk; c0 += MARCH_BS) {
      const int cnt = std::min(MARCH_BS, njk - c0);
      plc = std::max(plc, std::min(njk, c0 + cnt + 2 * d.nk) - std::max(0, c0 - 2 * d.nk));
      psy = std::max(psy, std::min(njk, c0 + cnt + d.nk) - std::max(0, c0 - d.nk));
      psz = std::max(psz, std::min(njk, c0 + cnt + 1) - std::max(0, c0 - 1));
      const int rows = (c0 + cnt - 1) / d.nk - c0 / d.nk;
      plf = std::max(plf, (cnt + d.nk) + (cnt + 1 + rows));
    }
    // thin patches (the reference skips the faces of an axis with <= 2 cells) and
    // patches whose staged planes would not fit the LDS budget use the cell kernel
    const size_t need = (size_t)(10 * plc + 5 * (plf + psy + psz)) * (ctx->prec == DRN_F32 ? 4 : 8);
    if (d.ni < 3 || d.nj < 3 || d.nk < 3 || d.nk > 1000 || njk > (1 << 20) || need > 120 * 1024) {
      ctx->thin.push_back(p);
      for (int t = ctx->tile_first[p]; t < ctx->tile_first[p + 1]; ++t) ctx->thin_tiles.push_back(ctx->tiles[t]);
      ctx->wt_first[p] = (int)ctx->wt_pp.size();
      continue;
    }
    lc = std::max(lc, plc);
    lf = std::max(lf, plf);
    lsy = std::max(lsy, psy);
    lsz = std::max(lsz, psz);
    march_items(d, p, tb.irange, chunk_of(ctx, ctx->tiling), nat);
    march_items(d, p, tp.irange, chunk_of(ctx, ctx->tiling_pp), ctx->mt_pp);
    ctx->wt_first[p] = (int)ctx->wt_pp.size();
    if (tb.kind == K_WTILE) wtile_items(d, p, tb.irange, wnat);
    if (tp.kind == K_WTILE) wtile_items(d, p, tp.irange, ctx->wt_pp);
  }
  ctx->mt_first[np] = (int)ctx->mt_pp.size();
  ctx->wt_first[np] = (int)ctx->wt_pp.size();
  ctx->lds_cells = lc;
  ctx->lds_faces = lf;
  ctx->lds_sy = lsy;
  ctx->lds_sz = lsz;
  ctx->n_mt = (int)nat.size();
  std::vector<TileMarch> launch = nat;
  if (tb.xcd && ctx->n_mt > 0) {
    // workgroup b runs on XCD b % 8: give each XCD one contiguous segment of the natural list
    const int n = ctx->n_mt, seg = (n + 7) / 8;
    std::vector<TileMarch> l;
    l.reserve(n);
    for (int r = 0; r < seg; ++r)
      for (int x = 0; x < 8; ++x)
        if (x * seg + r < n) l.push_back(nat[x * seg + r]);
    launch.swap(l);
  }
  // wave-tile launch list: blocks of 4 items, segmented per XCD like the march list
  ctx->n_wt = (int)wnat.size();
  std::vector<TileWave> wlaunch = wnat;
  if (tb.kind == K_WTILE && tb.xcd && ctx->n_wt > 0) {
    const int nb = (ctx->n_wt + 3) / 4, seg = (nb + 7) / 8;
    std::vector<TileWave> l;
    l.reserve(ctx->n_wt);
    for (int r = 0; r < seg; ++r)
      for (int x = 0; x < 8; ++x) {
        const int b = x * seg + r;
        if (b < nb)
          for (int q = 4 * b; q < std::min(4 * b + 4, ctx->n_wt); ++q) l.push_back(wnat[q]);
      }
    wlaunch.swap(l);
  }
  dfree(ctx->d_wt_pp);
  dfree(ctx->d_wt_launch);
  ctx->d_wt_pp = ctx->d_wt_launch = nullptr;
  HIPCHK(ctx, dalloc(&ctx->d_wt_pp, ctx->wt_pp.size()));
  HIPCHK(ctx, dalloc(&ctx->d_wt_launch, wlaunch.size()));
  if (!ctx->wt_pp.empty())
    HIPCHK(ctx, hipMemcpy(ctx->d_wt_pp, ctx->wt_pp.data(), ctx->wt_pp.size() * sizeof(TileWave), hipMemcpyHostToDevice));
  if (!wlaunch.empty())
    HIPCHK(ctx, hipMemcpy(ctx->d_wt_launch, wlaunch.data(), wlaunch.size() * sizeof(TileWave), hipMemcpyHostToDevice));
  dfree(ctx->d_mt_pp);
  dfree(ctx->d_mt_launch);
  dfree(ctx->d_thin_tiles);
  ctx->d_mt_pp = ctx->d_mt_launch = nullptr;
  ctx->d_thin_tiles = nullptr;
  HIPCHK(ctx, dalloc(&ctx->d_mt_pp, ctx->mt_pp.size()));
  HIPCHK(ctx, dalloc(&ctx->d_mt_launch, launch.size()));
  HIPCHK(ctx, dalloc(&ctx->d_thin_tiles, ctx->thin_tiles.size()));
  if (!ctx->mt_pp.empty())
    HIPCHK(ctx, hipMemcpy(ctx->d_mt_pp, ctx->mt_pp.data(), ctx->mt_pp.size() * sizeof(TileMarch), hipMemcpyHostToDevice));
  if (!launch.empty())
    HIPCHK(ctx, hipMemcpy(ctx->d_mt_launch, launch.data(), launch.size() * sizeof(TileMarch), hipMemcpyHostToDevice));
  if (!ctx->thin_tiles.empty())
    HIPCHK(ctx, hipMemcpy(ctx->d_thin_tiles, ctx->thin_tiles.data(), ctx->thin_tiles.size() * sizeof(TileCell),
                          hipMemcpyHostToDevice));
  return DRN_OK;
}

template <typename T>
static void launch_exchange_t(drn_ctx* ctx, const drn_ctx::ExchangeLists& x, T* U) {
  for (const auto& c : x.single) {
    if (c.n == 0) continue;
    hipLaunchKernelGGL(k_exchange_single<T>, dim3((c.n + 255) / 256), dim3(256), 0, ctx->stream, c.n, c.stride, c.rec,
                       c.idx, (const T*)c.w, U, ctx->vstride);
  }
  if (x.multi_n > 0 && x.m_et) {
    hipLaunchKernelGGL((k_exchange_multi<T, true>), dim3((x.multi_n + 255) / 256), dim3(256), 0, ctx->stream,
                       x.multi_n, x.m_rec, x.m_cnt, x.m_tbase, x.m_ent, (const T*)x.m_w, U, ctx->vstride, x.m_et,
                       (const T*)x.m_turn);
  } else if (x.multi_n > 0) {
    hipLaunchKernelGGL((k_exchange_multi<T, false>), dim3((x.multi_n + 255) / 256), dim3(256), 0, ctx->stream,
                       x.multi_n, x.m_rec, x.m_cnt, x.m_tbase, x.m_ent, (const T*)x.m_w, U, ctx->vstride, nullptr,
                       nullptr);
  }
}

// partial: the stage just launched filled the fused z-face receivers itself
static int do_exchange(drn_ctx* ctx, int pbuf, bool partial = false) {
  if (!ctx->have_exchange) return DRN_OK;
  const drn_ctx::ExchangeLists& x = partial ? ctx->xpart : ctx->xfull;
  hipEvent_t stop;
  timer_begin(ctx, 1, &stop);
  if (ctx->prec == DRN_F32)
    launch_exchange_t<float>(ctx, x, (float*)ctx->buf[pbuf]);
  else
    launch_exchange_t<double>(ctx, x, (double*)ctx->buf[pbuf]);
  timer_end(ctx, stop);
  HIPCHK(ctx, hipGetLastError());
  // after a stage that filled the fused receivers, the partial lists complete the exchange
  ctx->exch_buf = pbuf;
  return DRN_OK;
}

template <typename T>
static void launch_ls_t(drn_ctx* ctx, T* U, T relax) {
  const uint32_t n = ctx->ls_n;
  hipLaunchKernelGGL(k_ls_access<T>, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, ctx->ls_donors,
                     (const T*)ctx->ls_w, (const T*)U, ctx->vstride, (T*)ctx->ls_acc);
  if (ctx->ls_op == DRN_LS_SWALL)
    hipLaunchKernelGGL(k_ls_operate_swall<T>, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, ctx->ls_cells,
                       (const T*)ctx->ls_acc, U, ctx->vstride, relax);
  else
    hipLaunchKernelGGL(k_ls_operate<T>, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, ctx->ls_cells,
                       (const T*)ctx->ls_nrm, (const T*)ctx->ls_acc, U, ctx->vstride, relax);
}

static int do_ls_fill(drn_ctx* ctx, int pbuf, double relax) {
  if (ctx->ls_n == 0) return DRN_OK;
  wrote(ctx, pbuf);
  hipEvent_t stop;
  timer_begin(ctx, 2, &stop);
  if (ctx->prec == DRN_F32)
    launch_ls_t<float>(ctx, (float*)ctx->buf[pbuf], (float)relax);
  else
    launch_ls_t<double>(ctx, (double*)ctx->buf[pbuf], relax);
  timer_end(ctx, stop);
  HIPCHK(ctx, hipGetLastError());
  return DRN_OK;
}

// RungeKutta::operator() (rungekutta.cpp:27-42) on physical buffers: the state
// buffer X is exchanged (when needed), kept as U^n (base) for every stage, the
// stages alternate between the two other buffers, and the last stage output
// becomes logical field 0 while X becomes field 1 (the reference's copy of
// field 0 into field 1 at the step start).
static int do_rk_step(drn_ctx* ctx, double dt, int n_alpha, const double* alpha, int launch) {
  const bool post = ctx->ls_n > 0;
  const int X = ctx->map[0];
  const int Y = ctx->map[1] == X ? ctx->map[2] : ctx->map[1];
  int Z = 0;
  while (Z == X || Z == Y) ++Z;
  int rc;
  if (post || ctx->exch_buf != X) {
    if ((rc = do_exchange(ctx, X))) return rc;
  }
  int cur = X;
  // the stage fills the fused z-face receivers, the exchange after it the rest
  const bool push = push_active(ctx, launch);
  for (int s = 0; s < n_alpha; ++s) {
    if (s > 0 && post) {
      if ((rc = do_exchange(ctx, cur))) return rc;
    }
    const int dst = (s % 2 == 0) ? Y : Z;
    if ((rc = do_stage(ctx, cur, X, dst, alpha[s] * dt, launch, push, true))) return rc;
    if ((rc = do_exchange(ctx, dst, push))) return rc;
    if (post) {
      if ((rc = do_ls_fill(ctx, dst, 0.5))) return rc;
    }
    cur = dst;
  }
  ctx->map[0] = cur;
  ctx->map[1] = X;
  ctx->map[2] = (cur == Y) ? Z : Y;
  return DRN_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

// The functions below were declared extern "C" by drnhip.h and keep C linkage.

// Partial sums of drn_core_sums_noalloc: 5 doubles per plane chunk of the
// stage work list, at most kScratchChunks chunks (40 MiB on every device the
// library runs on, held until the process ends).  In host code `scratch` names
// a host-side stand-in: kernels and copies take the address that
// hipGetSymbolAddress gives for the current device, kept per device below.
constexpr size_t kScratchChunks = (size_t)1 << 20;
static __device__ __align__(256) double scratch[5 * kScratchChunks];
static std::mutex g_scratch_mu;
static std::map<int, double*> g_scratch_dev;

// the device address of `scratch` on ctx->device (current), looked up once per device
static int scratch_address(drn_ctx* ctx) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_dev.find(ctx->device);
  if (it == g_scratch_dev.end()) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return fail(ctx, DRN_E_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    }
    it = g_scratch_dev.emplace(ctx->device, (double*)p).first;
  }
  ctx->scratch = it->second;
  return DRN_OK;
}

int drn_ctx_create(int device, int precision, drn_ctx** out) {
  if (!out) return DRN_E_ARG;
  *out = nullptr;
  if (precision != DRN_F32 && precision != DRN_F64) return DRN_E_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return DRN_E_HIP;
  drn_ctx* ctx = new drn_ctx();
  ctx->device = device;
  ctx->prec = precision;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipDeviceGetAttribute(&ctx->n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      ctx->n_cu <= 0) {
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return DRN_E_HIP;
  }
  if (int rc = scratch_address(ctx)) {
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return rc;
  }
  *out = ctx;
  return DRN_OK;
}

static void free_lists(drn_ctx::ExchangeLists& x) {
  for (auto& c : x.single) {
    dfree(c.rec);
    dfree(c.idx);
    dfree(c.w);
  }
  x.single.clear();
  dfree(x.m_rec);
  dfree(x.m_cnt);
  dfree(x.m_tbase);
  dfree(x.m_ent);
  dfree(x.m_w);
  dfree(x.m_et);
  dfree(x.m_turn);
  x.m_rec = x.m_cnt = x.m_tbase = x.m_ent = x.m_et = nullptr;
  x.m_w = x.m_turn = nullptr;
  x.multi_n = 0;
}

static void free_exchange(drn_ctx* ctx) {
  free_lists(ctx->xfull);
  free_lists(ctx->xpart);
  ctx->have_exchange = false;
  ctx->n_fused = 0;
}

static void free_ls(drn_ctx* ctx) {
  dfree(ctx->ls_cells);
  dfree(ctx->ls_donors);
  dfree(ctx->ls_w);
  dfree(ctx->ls_nrm);
  dfree(ctx->ls_acc);
  ctx->ls_cells = ctx->ls_donors = nullptr;
  ctx->ls_w = ctx->ls_nrm = ctx->ls_acc = nullptr;
  ctx->ls_n = 0;
}

void drn_ctx_destroy(drn_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  drop_graph(ctx);
  free_exchange(ctx);
  free_ls(ctx);
  for (int b = 0; b < NBUF; ++b) dfree(ctx->buf[b]);
  dfree(ctx->d_pd);
  dfree(ctx->d_pd_rk);
  dfree(ctx->d_flags);
  dfree(ctx->d_tiles);
  dfree(ctx->d_tiles_xcd);
  dfree(ctx->d_tiles_xcd_pp);
  dfree(ctx->d_tz);
  dfree(ctx->d_tz_xcd);
  dfree(ctx->d_tz_chain);
  dfree(ctx->d_tz_xcd_pp);
  dfree(ctx->d_tyz);
  dfree(ctx->d_tyz_xcd);
  dfree(ctx->d_tyz_xcd_pp);
  dfree(ctx->d_tx2);
  dfree(ctx->d_tx2_xcd);
  dfree(ctx->d_tx2_xcd_pp);
  for (auto& L : ctx->xs) {
    dfree(L.d);
    dfree(L.d_xcd);
    dfree(L.d_xcd_pp);
    dfree(L.d_chain);
    L.d = L.d_xcd = L.d_xcd_pp = L.d_chain = nullptr;
  }
  dfree(ctx->d_mt_pp);
  dfree(ctx->d_mt_launch);
  dfree(ctx->d_thin_tiles);
  dfree(ctx->d_wt_pp);
  dfree(ctx->d_wt_launch);
  dfree(ctx->d_part);
  dfree(ctx->d_out);
  dfree(ctx->d_post);
  for (auto& s : ctx->timers) {
    for (auto& e : s.used) {
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
    for (auto& e : s.pool) {
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
  }
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* drn_last_error(const drn_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int drn_precision(const drn_ctx* ctx) { return ctx ? ctx->prec : DRN_E_ARG; }

int drn_set_operator(drn_ctx* ctx, int recon, int limiter) {
  if (!ctx) return DRN_E_ARG;
  if (recon != DRN_RECON_UPWIND1 && recon != DRN_RECON_UPWIND2) return fail(ctx, DRN_E_ARG, "unknown reconstruction");
  if (limiter < DRN_LIM_VANALBADA || limiter > DRN_LIM_SECONDORDER) return fail(ctx, DRN_E_ARG, "unknown limiter");
  ctx->recon = recon;
  ctx->lim = limiter;
  if (ctx->laid_out) {
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    int rc = build_march(ctx);
    if (rc) return rc;
  }
  drop_graph(ctx);
  return DRN_OK;
}

int drn_set_viscous(drn_ctx* ctx, int on) {
  if (!ctx) return DRN_E_ARG;
  if (on != 0 && on != 1) return fail(ctx, DRN_E_ARG, "viscous switch must be 0 or 1");
  if (ctx->laid_out) HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->visc = on != 0;
  // the kernel kinds of the tiling choices change with the switch
  if (ctx->laid_out) {
    int rc = build_march(ctx);
    if (rc) return rc;
  }
  drop_graph(ctx);
  return DRN_OK;
}

int drn_patch_add(drn_ctx* ctx, const drn_patch_desc* d, int* id) {
  if (!ctx || !d) return DRN_E_ARG;
  if (ctx->laid_out) return fail(ctx, DRN_E_STATE, "patches must be added before drn_layout");
  if (d->ni < 1 || d->nj < 1 || d->nk < 1) return fail(ctx, DRN_E_ARG, "patch sizes must be positive");
  if (d->ni > 32767 || d->nj > 32767 || d->nk > 32767) return fail(ctx, DRN_E_ARG, "patch side above 32767 cells");
  if ((uint64_t)d->ni * d->nj * d->nk > (1ull << 31)) return fail(ctx, DRN_E_ARG, "patch too large");
  for (int f = 0; f < 6; ++f) {
    if (d->bc[f] < DRN_BC_SLIP || d->bc[f] > DRN_BC_NONE) return fail(ctx, DRN_E_ARG, "unknown face functor");
    if (d->seek[f] < 0) return fail(ctx, DRN_E_ARG, "negative seek layer count");
  }
  if (!(d->dx > 0 && d->dy > 0 && d->dz > 0)) return fail(ctx, DRN_E_ARG, "spacing must be positive");
  ctx->descs.push_back(*d);
  if (id) *id = (int)ctx->descs.size() - 1;
  return DRN_OK;
}

int drn_num_patches(const drn_ctx* ctx) { return ctx ? (int)ctx->descs.size() : DRN_E_ARG; }

size_t drn_patch_cells(const drn_ctx* ctx, int id) {
  if (!ctx || id < 0 || id >= (int)ctx->descs.size()) return 0;
  const auto& d = ctx->descs[id];
  return (size_t)d.ni * d.nj * d.nk;
}

int drn_layout(drn_ctx* ctx) {
  if (!ctx) return DRN_E_ARG;
  if (ctx->laid_out) return fail(ctx, DRN_E_STATE, "layout already done");
  if (ctx->descs.empty()) return fail(ctx, DRN_E_STATE, "no patches");
  {
    // the kernels address a variable array with 32-bit byte offsets: check the
    // padded size before building any list or allocating anything
    uint64_t cells = 0;
    for (const auto& d : ctx->descs)
      cells += ((uint64_t)d.ni * d.nj * d.nk + ALIGN_CELLS - 1) / ALIGN_CELLS * ALIGN_CELLS;
    if (cells * ctx->esize() >= (1ull << 32))
      return fail(ctx, DRN_E_ARG, "grid exceeds 2^32 bytes per variable array (2^30 FP32 / 2^29 FP64 cells)");
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const int np = (int)ctx->descs.size();
  std::vector<PatchDev> pd(np);
  uint64_t off = 0;
  ctx->offs.resize(np);
  ctx->tile_first.assign(np + 1, 0);
  ctx->tiles.clear();
  for (int p = 0; p < np; ++p) {
    const auto& d = ctx->descs[p];
    PatchDev& P = pd[p];
    memset(&P, 0, sizeof(P));
    P.ni = d.ni;
    P.nj = d.nj;
    P.nk = d.nk;
    P.njk = d.nj * d.nk;
    P.off = (uint32_t)off;
    P.dx = d.dx;
    P.dy = d.dy;
    P.dz = d.dz;
    for (int f = 0; f < 6; ++f) {
      P.bc[f] = d.bc[f];
      P.seek[f] = d.seek[f];
      for (int q = 0; q < 16; ++q) P.par[f][q] = d.bc_par[f][q];
    }
    ctx->offs[p] = off;
    const uint64_t n = (uint64_t)d.ni * d.nj * d.nk;
    off += (n + ALIGN_CELLS - 1) / ALIGN_CELLS * ALIGN_CELLS;
    ctx->tile_first[p] = (int)ctx->tiles.size();
    const int njk = d.nj * d.nk;
    for (int i = 0; i < d.ni; ++i)
      for (int c2 = 0; c2 < njk; c2 += 256) ctx->tiles.push_back(TileCell{p, i, c2, 0});
  }
  ctx->tile_first[np] = (int)ctx->tiles.size();
  ctx->tz.clear();
  ctx->tz_first.assign(np + 1, 0);
  for (int p = 0; p < np; ++p) {
    const auto& d = ctx->descs[p];
    ctx->tz_first[p] = (int)ctx->tz.size();
    const int njk = d.nj * d.nk;
    for (int i = 0; i < d.ni; ++i)
      for (int c2 = 0; c2 < njk; c2 += 252) ctx->tz.push_back(TileCell{p, i, c2, 0});
  }
  ctx->tz_first[np] = (int)ctx->tz.size();
  ctx->tyz.clear();
  ctx->tyz_first.assign(np + 1, 0);
  for (int p = 0; p < np; ++p) {
    const auto& d = ctx->descs[p];
    ctx->tyz_first[p] = (int)ctx->tyz.size();
    const int njk = d.nj * d.nk;
    for (int i = 0; i < d.ni; ++i)
      for (int c2 = 0; c2 < njk; c2 += 504) ctx->tyz.push_back(TileCell{p, i, c2, 0});
  }
  ctx->tyz_first[np] = (int)ctx->tyz.size();
  for (int q = 0; q < 2; ++q) {
    const int XP = q ? 4 : 2, step = 252 / XP;
    auto& L = ctx->xs[q];
    L.nat.clear();
    L.first.assign(np + 1, 0);
    for (int p = 0; p < np; ++p) {
      const auto& d = ctx->descs[p];
      L.first[p] = (int)L.nat.size();
      const int njk = d.nj * d.nk;
      for (int i = 0; i < d.ni; i += XP)
        for (int c2 = 0; c2 < njk; c2 += step) L.nat.push_back(TileCell{p, i, c2, 0});
    }
    L.first[np] = (int)L.nat.size();
  }
  ctx->tx2.clear();
  ctx->tx2_first.assign(np + 1, 0);
  for (int p = 0; p < np; ++p) {
    const auto& d = ctx->descs[p];
    ctx->tx2_first[p] = (int)ctx->tx2.size();
    const int njk = d.nj * d.nk;
    for (int i = 0; i < d.ni; i += 2)
      for (int c2 = 0; c2 < njk; c2 += 252) ctx->tx2.push_back(TileCell{p, i, c2, 0});
  }
  ctx->tx2_first[np] = (int)ctx->tx2.size();
  ctx->vstride = off;
  for (int b = 0; b < NBUF; ++b) {
    if (hipMalloc(&ctx->buf[b], 5 * off * ctx->esize()) != hipSuccess) return fail(ctx, DRN_E_NOMEM, "state buffer allocation failed");
    HIPCHK(ctx, hipMemset(ctx->buf[b], 0, 5 * off * ctx->esize()));
  }
  HIPCHK(ctx, dalloc(&ctx->d_pd, np));
  HIPCHK(ctx, hipMemcpy(ctx->d_pd, pd.data(), np * sizeof(PatchDev), hipMemcpyHostToDevice));
  ctx->pd_host = pd;
  HIPCHK(ctx, dalloc(&ctx->d_tiles, ctx->tiles.size()));
  HIPCHK(ctx, hipMemcpy(ctx->d_tiles, ctx->tiles.data(), ctx->tiles.size() * sizeof(TileCell), hipMemcpyHostToDevice));
  {
    // XCD-segmented orders of the cell list: workgroup b runs on XCD b % 8, so
    // launch slot r*8 + x takes item x*seg + r of the natural list and each XCD
    // walks one contiguous run of planes (its L2 then holds the i+-2 halo planes
    // its own workgroups need)
    auto segment = [](const TileCell* nat, int n, TileCell* out) {
      const int seg = (n + 7) / 8;
      int m = 0;
      for (int r = 0; r < seg; ++r)
        for (int x = 0; x < 8; ++x)
          if (x * seg + r < n) out[m++] = nat[x * seg + r];
    };
    const int nt = (int)ctx->tiles.size();
    std::vector<TileCell> all(nt), pp(nt);
    segment(ctx->tiles.data(), nt, all.data());
    for (int p = 0; p < np; ++p)
      segment(ctx->tiles.data() + ctx->tile_first[p], ctx->tile_first[p + 1] - ctx->tile_first[p],
              pp.data() + ctx->tile_first[p]);
    HIPCHK(ctx, dalloc(&ctx->d_tiles_xcd, nt));
    HIPCHK(ctx, dalloc(&ctx->d_tiles_xcd_pp, nt));
    HIPCHK(ctx, hipMemcpy(ctx->d_tiles_xcd, all.data(), nt * sizeof(TileCell), hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(ctx->d_tiles_xcd_pp, pp.data(), nt * sizeof(TileCell), hipMemcpyHostToDevice));
    const int nz = (int)ctx->tz.size();
    std::vector<TileCell> zall(nz), zpp(nz);
    segment(ctx->tz.data(), nz, zall.data());
    for (int p = 0; p < np; ++p)
      segment(ctx->tz.data() + ctx->tz_first[p], ctx->tz_first[p + 1] - ctx->tz_first[p], zpp.data() + ctx->tz_first[p]);
    HIPCHK(ctx, dalloc(&ctx->d_tz, nz));
    HIPCHK(ctx, dalloc(&ctx->d_tz_xcd, nz));
    HIPCHK(ctx, dalloc(&ctx->d_tz_xcd_pp, nz));
    HIPCHK(ctx, hipMemcpy(ctx->d_tz, ctx->tz.data(), nz * sizeof(TileCell), hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(ctx->d_tz_xcd, zall.data(), nz * sizeof(TileCell), hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(ctx->d_tz_xcd_pp, zpp.data(), nz * sizeof(TileCell), hipMemcpyHostToDevice));
    const int ny = (int)ctx->tyz.size();
    std::vector<TileCell> yall(ny), ypp(ny);
    segment(ctx->tyz.data(), ny, yall.data());
    for (int p = 0; p < np; ++p)
      segment(ctx->tyz.data() + ctx->tyz_first[p], ctx->tyz_first[p + 1] - ctx->tyz_first[p],
              ypp.data() + ctx->tyz_first[p]);
    HIPCHK(ctx, dalloc(&ctx->d_tyz, ny));
    HIPCHK(ctx, dalloc(&ctx->d_tyz_xcd, ny));
    HIPCHK(ctx, dalloc(&ctx->d_tyz_xcd_pp, ny));
    HIPCHK(ctx, hipMemcpy(ctx->d_tyz, ctx->tyz.data(), ny * sizeof(TileCell), hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(ctx->d_tyz_xcd, yall.data(), ny * sizeof(TileCell), hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(ctx->d_tyz_xcd_pp, ypp.data(), ny * sizeof(TileCell), hipMemcpyHostToDevice));
    for (auto& L : ctx->xs) {
      const int n = (int)L.nat.size();
      std::vector<TileCell> all(n), pp(n);
      segment(L.nat.data(), n, all.data());
      for (int p = 0; p < np; ++p) segment(L.nat.data() + L.first[p], L.first[p + 1] - L.first[p], pp.data() + L.first[p]);
      HIPCHK(ctx, dalloc(&L.d, n));
      HIPCHK(ctx, dalloc(&L.d_xcd, n));
      HIPCHK(ctx, dalloc(&L.d_xcd_pp, n));
      HIPCHK(ctx, hipMemcpy(L.d, L.nat.data(), n * sizeof(TileCell), hipMemcpyHostToDevice));
      HIPCHK(ctx, hipMemcpy(L.d_xcd, all.data(), n * sizeof(TileCell), hipMemcpyHostToDevice));
      HIPCHK(ctx, hipMemcpy(L.d_xcd_pp, pp.data(), n * sizeof(TileCell), hipMemcpyHostToDevice));
    }
    const int n2 = (int)ctx->tx2.size();
    std::vector<TileCell> xall(n2), xpp(n2);
    segment(ctx->tx2.data(), n2, xall.data());
    for (int p = 0; p < np; ++p)
      segment(ctx->tx2.data() + ctx->tx2_first[p], ctx->tx2_first[p + 1] - ctx->tx2_first[p],
              xpp.data() + ctx->tx2_first[p]);
    HIPCHK(ctx, dalloc(&ctx->d_tx2, n2));
    HIPCHK(ctx, dalloc(&ctx->d_tx2_xcd, n2));
    HIPCHK(ctx, dalloc(&ctx->d_tx2_xcd_pp, n2));
    HIPCHK(ctx, hipMemcpy(ctx->d_tx2, ctx->tx2.data(), n2 * sizeof(TileCell), hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(ctx->d_tx2_xcd, xall.data(), n2 * sizeof(TileCell), hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(ctx->d_tx2_xcd_pp, xpp.data(), n2 * sizeof(TileCell), hipMemcpyHostToDevice));
  }
  HIPCHK(ctx, dalloc(&ctx->d_part, ctx->tiles.size() * 5));
  HIPCHK(ctx, dalloc(&ctx->d_out, 8));
  int rc = build_march(ctx);
  if (rc) return rc;
  ctx->laid_out = true;
  return DRN_OK;
}

static int check_field(drn_ctx* ctx, int field) {
  if (!ctx->laid_out) return fail(ctx, DRN_E_STATE, "drn_layout not called");
  if (field < 0 || field >= NBUF) return fail(ctx, DRN_E_ARG, "field must be 0, 1 or 2");
  return DRN_OK;
}

int drn_patch_upload(drn_ctx* ctx, int id, int field, const void* host) {
  if (!ctx || !host) return DRN_E_ARG;
  int rc;
  if ((rc = check_field(ctx, field))) return rc;
  if (id < 0 || id >= (int)ctx->descs.size()) return fail(ctx, DRN_E_ARG, "patch id out of range");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t n = drn_patch_cells(ctx, id), es = ctx->esize();
  wrote(ctx, ctx->map[field]);
  char* b = (char*)ctx->buf[ctx->map[field]];
  for (int v = 0; v < 5; ++v)
    HIPCHK(ctx, hipMemcpyAsync(b + (v * ctx->vstride + ctx->offs[id]) * es, (const char*)host + v * n * es, n * es,
                               hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return DRN_OK;
}

int drn_patch_download(drn_ctx* ctx, int id, int field, void* host) {
  if (!ctx || !host) return DRN_E_ARG;
  int rc;
  if ((rc = check_field(ctx, field))) return rc;
  if (id < 0 || id >= (int)ctx->descs.size()) return fail(ctx, DRN_E_ARG, "patch id out of range");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t n = drn_patch_cells(ctx, id), es = ctx->esize();
  const char* b = (const char*)ctx->buf[ctx->map[field]];
  for (int v = 0; v < 5; ++v)
    HIPCHK(ctx, hipMemcpyAsync((char*)host + v * n * es, b + (v * ctx->vstride + ctx->offs[id]) * es, n * es,
                               hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return DRN_OK;
}

int drn_copy_field(drn_ctx* ctx, int src, int dst) {
  if (!ctx) return DRN_E_ARG;
  int rc;
  if ((rc = check_field(ctx, src)) || (rc = check_field(ctx, dst))) return rc;
  if (src == dst) return DRN_OK;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  // buffers hold 5*vstride values, vstride a multiple of 64 cells: a whole number of 16-byte units
  const uint64_t n16 = 5 * ctx->vstride * ctx->esize() / 16;
  wrote(ctx, ctx->map[dst]);
  hipEvent_t stop;
  timer_begin(ctx, 3, &stop);
  hipLaunchKernelGGL(k_copy, dim3(2048), dim3(256), 0, ctx->stream, n16, (const uint4*)ctx->buf[ctx->map[src]],
                     (uint4*)ctx->buf[ctx->map[dst]]);
  timer_end(ctx, stop);
  HIPCHK(ctx, hipGetLastError());
  return DRN_OK;
}

int drn_swap_fields(drn_ctx* ctx, int a, int b) {
  if (!ctx) return DRN_E_ARG;
  int rc;
  if ((rc = check_field(ctx, a)) || (rc = check_field(ctx, b))) return rc;
  std::swap(ctx->map[a], ctx->map[b]);
  return DRN_OK;
}

int drn_set_cell_flags(drn_ctx* ctx, int patch, const uint8_t* flags) {
  if (!ctx) return DRN_E_ARG;
  if (!ctx->laid_out) return fail(ctx, DRN_E_STATE, "drn_layout not called");
  const int np = (int)ctx->descs.size();
  if (patch < 0 || patch >= np) return fail(ctx, DRN_E_ARG, "patch id out of range");
  const size_t n = drn_patch_cells(ctx, patch);
  bool any = false;
  if (flags)
    for (size_t l = 0; l < n; ++l) {
      if (flags[l] & ~(uint8_t)(DRN_CELL_INACTIVE | DRN_CELL_INSIDE))
        return fail(ctx, DRN_E_ARG, "unknown cell flag bit");
      any = any || flags[l] != 0;
    }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (!ctx->d_flags) {
    if (!any) return DRN_OK;
    HIPCHK(ctx, dalloc(&ctx->d_flags, ctx->vstride));
    HIPCHK(ctx, hipMemset(ctx->d_flags, 0, ctx->vstride));
    ctx->flags_any.assign(np, 0);
  }
  if (any)
    HIPCHK(ctx, hipMemcpy(ctx->d_flags + ctx->offs[patch], flags, n, hipMemcpyHostToDevice));
  else
    HIPCHK(ctx, hipMemset(ctx->d_flags + ctx->offs[patch], 0, n));
  ctx->flags_any[patch] = any ? 1 : 0;
  bool on = false;
  for (uint8_t a : ctx->flags_any) on = on || a;
  if (on != ctx->flags_on) {
    ctx->flags_on = on;
    drop_graph(ctx);
    int rc = build_march(ctx);
    if (rc) return rc;
  }
  return DRN_OK;
}

int drn_stage(drn_ctx* ctx, int src, int base, int dst, double factor, int launch) {
  if (!ctx) return DRN_E_ARG;
  int rc;
  if ((rc = check_field(ctx, src)) || (rc = check_field(ctx, base)) || (rc = check_field(ctx, dst))) return rc;
  if (src == dst) return fail(ctx, DRN_E_ARG, "drn_stage: dst must differ from src");
  if (launch != DRN_LAUNCH_BATCHED && launch != DRN_LAUNCH_PER_PATCH) return fail(ctx, DRN_E_ARG, "unknown launch form");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  return do_stage(ctx, ctx->map[src], ctx->map[base], ctx->map[dst], factor, launch);
}

int drn_exchange_set(drn_ctx* ctx, int recv_id, int n_donors, const int32_t* donor_patch, const int64_t* donor_info,
                     int64_t n_unique, const int64_t* rec_unique, int64_t n_rec_concat, const int64_t* rec_concat,
                     int64_t n_wi, const int64_t* donor_index, const double* donor_weight) {
  if (!ctx) return DRN_E_ARG;
  const int np = (int)ctx->descs.size();
  if (recv_id < 0 || recv_id >= np) return fail(ctx, DRN_E_ARG, "receiver patch id out of range");
  if (n_donors < 0 || n_unique < 0 || n_rec_concat < 0 || n_wi < 0) return fail(ctx, DRN_E_ARG, "negative count");
  if ((n_donors > 0 && (!donor_patch || !donor_info)) || (n_unique > 0 && !rec_unique) ||
      (n_rec_concat > 0 && !rec_concat) || (n_wi > 0 && (!donor_index || !donor_weight)))
    return fail(ctx, DRN_E_ARG, "null table pointer with a positive count");
  const int64_t nrecv_cells = (int64_t)drn_patch_cells(ctx, recv_id);
  ExchangeTables t;
  t.donor_patch.assign(donor_patch, donor_patch + n_donors);
  t.donor_info.assign(donor_info, donor_info + 4 * (size_t)n_donors);
  t.rec_unique.assign(rec_unique, rec_unique + n_unique);
  t.rec_concat.assign(rec_concat, rec_concat + n_rec_concat);
  t.donor_index.assign(donor_index, donor_index + n_wi);
  t.donor_weight.assign(donor_weight, donor_weight + n_wi);
  for (int d = 0; d < n_donors; ++d) {
    const int dp = t.donor_patch[d];
    if (dp < 0 || dp >= np) return fail(ctx, DRN_E_RANGE, "donor patch id out of range");
    const int64_t nrec = t.donor_info[4 * d], stride = t.donor_info[4 * d + 1], r0 = t.donor_info[4 * d + 2],
                  w0 = t.donor_info[4 * d + 3];
    if (nrec < 0 || stride < 1 || stride > 64 || r0 < 0 || r0 + nrec > n_rec_concat || w0 < 0 ||
        w0 + nrec * stride > n_wi)
      return fail(ctx, DRN_E_RANGE, "donor table extents inconsistent");
    const int64_t ndc = (int64_t)drn_patch_cells(ctx, dp);
    for (int64_t a = w0; a < w0 + nrec * stride; ++a)
      if (t.donor_index[a] < 0 || t.donor_index[a] >= ndc) return fail(ctx, DRN_E_RANGE, "donor cell outside donor patch");
  }
  for (int64_t a = 0; a < n_rec_concat; ++a)
    if (t.rec_concat[a] < 0 || t.rec_concat[a] >= nrecv_cells) return fail(ctx, DRN_E_RANGE, "receiver cell outside patch");
  for (int64_t a = 0; a < n_unique; ++a)
    if (t.rec_unique[a] < 0 || t.rec_unique[a] >= nrecv_cells) return fail(ctx, DRN_E_RANGE, "receiver cell outside patch");
  ctx->xt[recv_id] = std::move(t);
  return DRN_OK;
}

int drn_exchange_set_turn(drn_ctx* ctx, int recv_id, int n_donors, const double* turn) {
  if (!ctx) return DRN_E_ARG;
  auto it = ctx->xt.find(recv_id);
  if (it == ctx->xt.end()) return fail(ctx, DRN_E_STATE, "drn_exchange_set_turn: no tables set for this receiver patch");
  if (n_donors != (int)it->second.donor_patch.size())
    return fail(ctx, DRN_E_ARG, "drn_exchange_set_turn: donor count differs from the tables");
  if (n_donors > 0 && !turn) return fail(ctx, DRN_E_ARG, "drn_exchange_set_turn: null matrix pointer");
  for (int64_t a = 0; a < 9 * (int64_t)n_donors; ++a)
    if (!std::isfinite(turn[a])) return fail(ctx, DRN_E_ARG, "drn_exchange_set_turn: non-finite matrix entry");
  it->second.turn.assign(turn, turn + 9 * (size_t)n_donors);
  return DRN_OK;
}

template <typename T>
static int upload_vec(drn_ctx* ctx, T** dptr, const std::vector<T>& h) {
  HIPCHK(ctx, dalloc(dptr, h.size()));
  if (!h.empty()) HIPCHK(ctx, hipMemcpy(*dptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return DRN_OK;
}

static int upload_weights(drn_ctx* ctx, void** dptr, const std::vector<double>& h) {
  if (ctx->prec == DRN_F32) {
    std::vector<float> f(h.begin(), h.end());
    float* p = nullptr;
    int rc = upload_vec(ctx, &p, f);
    *dptr = p;
    return rc;
  }
  double* p = nullptr;
  int rc = upload_vec(ctx, &p, h);
  *dptr = p;
  return rc;
}

// the patch that holds global cell g
static int patch_of(const drn_ctx* ctx, uint64_t g) {
  return (int)(std::upper_bound(ctx->offs.begin(), ctx->offs.end(), g) - ctx->offs.begin()) - 1;
}

// The z-shared list of tiling value xcd = 2: patches joined by fused z-face
// pushes form chains, and a chain's patches are walked plane by plane
// together (plane i of every patch of the chain, then plane i + 1), so that
// a donor's pushed rows and the receiver's own rows of the same lines are
// written close together, by one XCD; then segmented per XCD as the others.
static int build_chain_list(drn_ctx* ctx, const std::vector<std::pair<int, int>>& pairs) {
  dfree(ctx->d_tz_chain);
  ctx->d_tz_chain = nullptr;
  for (auto& L : ctx->xs) {
    dfree(L.d_chain);
    L.d_chain = nullptr;
  }
  const int np = (int)ctx->descs.size();
  if (pairs.empty() || np == 0) return DRN_OK;
  std::vector<int> up(np);
  for (int p = 0; p < np; ++p) up[p] = p;
  auto root = [&](int p) {
    while (up[p] != p) p = up[p] = up[up[p]];
    return p;
  };
  for (const auto& pr : pairs) {
    const int a = root(pr.first), b = root(pr.second);
    if (a != b) up[std::max(a, b)] = std::min(a, b);
  }
  std::vector<std::vector<int>> chains;
  std::vector<int> chain_of(np, -1);
  for (int p = 0; p < np; ++p) {
    const int r = root(p);
    if (chain_of[r] < 0) {
      chain_of[r] = (int)chains.size();
      chains.emplace_back();
    }
    chains[chain_of[r]].push_back(p);
  }
  // items of xp planes x `cells` cells, the planes of a chain's patches side by
  // side, then in XCD segments
  auto chain_list = [&](int xp, int cells, size_t n_expect, TileCell** dptr) -> int {
    std::vector<TileCell> nat;
    nat.reserve(n_expect);
    for (const auto& ch : chains) {
      int ni_max = 0;
      for (int p : ch) ni_max = std::max(ni_max, ctx->descs[p].ni);
      for (int i = 0; i < ni_max; i += xp)
        for (int p : ch) {
          const auto& d = ctx->descs[p];
          if (i >= d.ni) continue;
          const int njk = d.nj * d.nk;
          for (int c2 = 0; c2 < njk; c2 += cells) nat.push_back(TileCell{p, i, c2, 0});
        }
    }
    if (nat.size() != n_expect) return fail(ctx, DRN_E_STATE, "chain list does not cover the z-shared list");
    const int n = (int)nat.size(), seg = (n + 7) / 8;
    std::vector<TileCell> out;
    out.reserve(n);
    for (int r = 0; r < seg; ++r)
      for (int x = 0; x < 8; ++x)
        if (x * seg + r < n) out.push_back(nat[x * seg + r]);
    HIPCHK(ctx, dalloc(dptr, n));
    HIPCHK(ctx, hipMemcpy(*dptr, out.data(), n * sizeof(TileCell), hipMemcpyHostToDevice));
    return DRN_OK;
  };
  int rc;
  if ((rc = chain_list(1, 252, ctx->tz.size(), &ctx->d_tz_chain))) return rc;
  if ((rc = chain_list(2, 126, ctx->xs[0].nat.size(), &ctx->xs[0].d_chain))) return rc;
  return chain_list(4, 63, ctx->xs[1].nat.size(), &ctx->xs[1].d_chain);
}

int drn_exchange_commit(drn_ctx* ctx) {
  if (!ctx) return DRN_E_ARG;
  if (!ctx->laid_out) return fail(ctx, DRN_E_STATE, "drn_layout not called");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  drop_graph(ctx);  // a captured step holds the old lists
  free_exchange(ctx);
  for (PatchDev& P : ctx->pd_host) {
    P.npush = P.nrecv = 0;
    memset(P.push, 0, sizeof(P.push));
    memset(P.recv, 0, sizeof(P.recv));
  }
  // per unique receiver: its contribution groups in donor order
  struct Group {
    int stride;
    int turn = 0;  // 0 = aligned, t > 0 = matrix t - 1 of `turns`
    std::vector<uint32_t> idx;
    std::vector<double> w;
  };
  std::vector<double> turns;  // 9 per turned (receiver patch, donor) relation
  std::map<int, std::vector<std::pair<uint32_t, std::vector<Group>>>> single_by_stride;
  std::vector<std::pair<uint32_t, std::vector<Group>>> multi;
  std::vector<bool> fused(ctx->vstride, false);  // receivers the stage kernel fills when fused
  int64_t n_fused = 0;
  std::vector<std::pair<int, int>> push_pairs;    // (donor patch, receiving patch) of each fused side
  for (auto& kv : ctx->xt) {
    const int p = kv.first;
    const ExchangeTables& t = kv.second;
    std::map<int64_t, std::vector<Group>> groups;
    for (size_t d = 0; d < t.donor_patch.size(); ++d) {
      const int dp = t.donor_patch[d];
      const int64_t nrec = t.donor_info[4 * d], stride = t.donor_info[4 * d + 1], r0 = t.donor_info[4 * d + 2],
                    w0 = t.donor_info[4 * d + 3];
      int turn = 0;
      if (!t.turn.empty() && !is_identity_turn(&t.turn[9 * d])) {
        turns.insert(turns.end(), t.turn.begin() + 9 * d, t.turn.begin() + 9 * (d + 1));
        turn = (int)(turns.size() / 9);
      }
      for (int64_t ll = 0; ll < nrec; ++ll) {
        Group g;
        g.stride = (int)stride;
        g.turn = turn;
        for (int64_t s = 0; s < stride; ++s) {
          g.idx.push_back((uint32_t)(ctx->offs[dp] + (uint64_t)t.donor_index[w0 + ll * stride + s]));
          g.w.push_back(t.donor_weight[w0 + ll * stride + s]);
        }
        groups[t.rec_concat[r0 + ll]].push_back(std::move(g));
      }
    }
    // The z-face seek layers of the x/y core columns, one side at a time, are
    // filled by the donors' stage threads when every cell of the side is a
    // receiver with one unit-weight donor, all in one donor patch at one
    // (i, j, k) offset: the aligned patch-to-patch copy of a block grid.
    const drn_patch_desc& d = ctx->descs[p];
    const int64_t njk = (int64_t)d.nj * d.nk;
    const int i0 = d.seek[0], i1 = d.ni - d.seek[1], j0 = d.seek[2], j1 = d.nj - d.seek[3];
    for (int side = 0; side < 2; ++side) {
      const int k0 = side == 0 ? 0 : d.nk - d.seek[5];
      const int k1 = side == 0 ? d.seek[4] : d.nk;
      if (k1 <= k0 || i1 <= i0 || j1 <= j0 || k0 < 0) continue;
      int q = -1;
      int64_t oi = 0, oj = 0, ok = 0;
      auto aligned = [&]() {
        for (int i = i0; i < i1; ++i)
          for (int j = j0; j < j1; ++j)
            for (int k = k0; k < k1; ++k) {
              const auto it = groups.find((int64_t)i * njk + (int64_t)j * d.nk + k);
              if (it == groups.end() || it->second.size() != 1) return false;
              const Group& g = it->second[0];
              if (g.stride != 1 || g.w[0] != 1.0 || g.turn != 0) return false;
              const int dq = patch_of(ctx, g.idx[0]);
              const drn_patch_desc& e = ctx->descs[dq];
              const int64_t enjk = (int64_t)e.nj * e.nk, dl = (int64_t)(g.idx[0] - ctx->offs[dq]);
              const int64_t iq = dl / enjk, jq = (dl % enjk) / e.nk, kq = dl % e.nk;
              if (q < 0) {
                q = dq;
                oi = iq - i;
                oj = jq - j;
                ok = kq - k;
              } else if (dq != q || iq - i != oi || jq - j != oj || kq - k != ok) {
                return false;
              }
            }
        return q >= 0;
      };
      if (!aligned()) continue;
      PatchDev& Q = ctx->pd_host[q];
      PatchDev& P = ctx->pd_host[p];
      if (Q.npush >= 2 || P.nrecv >= 2) continue;
      const int sl = Q.npush++;
      Q.push[sl] = CellBox{(int32_t)(i0 + oi), (int32_t)(i1 + oi), (int32_t)(j0 + oj), (int32_t)(j1 + oj),
                           (int32_t)(k0 + ok), (int32_t)(k1 + ok)};
      Q.push_base[sl] = (int64_t)ctx->offs[p] - oi * njk - oj * d.nk - ok;
      Q.push_si[sl] = (uint32_t)njk;
      Q.push_sj[sl] = (uint32_t)d.nk;
      P.recv[P.nrecv++] = CellBox{i0, i1, j0, j1, k0, k1};
      push_pairs.push_back({q, p});
      for (int i = i0; i < i1; ++i)
        for (int j = j0; j < j1; ++j)
          for (int k = k0; k < k1; ++k) fused[ctx->offs[p] + (uint64_t)((int64_t)i * njk + (int64_t)j * d.nk + k)] = true;
      n_fused += (int64_t)(i1 - i0) * (j1 - j0) * (k1 - k0);
    }
    for (int64_t r : t.rec_unique) {
      auto it = groups.find(r);
      if (it == groups.end()) return fail(ctx, DRN_E_RANGE, "unique receiver without donor contribution");
      const uint32_t g = (uint32_t)(ctx->offs[p] + (uint64_t)r);
      // a receiver with a turned donor goes to the multi-donor list, whose
      // kernel turns the momentum of each donor sum
      bool turned = false;
      for (const Group& gr : it->second) turned = turned || gr.turn != 0;
      if (it->second.size() == 1 && !turned)
        single_by_stride[it->second[0].stride].push_back({g, std::move(it->second)});
      else
        multi.push_back({g, std::move(it->second)});
    }
  }
  // A donor cell that is also a receiver would make the result depend on the
  // order in which receivers are filled (the reference fills them patch after
  // patch, the exchange kernels all at once).  The reference's donor search
  // takes donors from the core of a patch only, so such tables are rejected.
  {
    std::vector<bool> is_recv(ctx->vstride, false);
    for (const auto& kv : single_by_stride)
      for (const auto& e : kv.second) is_recv[e.first] = true;
    for (const auto& e : multi) is_recv[e.first] = true;
    auto donor_is_recv = [&](const std::vector<std::pair<uint32_t, std::vector<Group>>>& list) {
      for (const auto& e : list)
        for (const auto& g : e.second)
          for (uint32_t dd : g.idx)
            if (is_recv[dd]) return true;
      return false;
    };
    bool clash = donor_is_recv(multi);
    for (const auto& kv : single_by_stride) clash = clash || donor_is_recv(kv.second);
    if (clash) return fail(ctx, DRN_E_RANGE, "a donor cell is also a receiver cell of the exchange");
  }
  // the full lists, and the lists without the fused receivers
  // Single-donor receivers of stride > 1 go to the multi-donor list: one
  // launch fewer, the same sum (0 + the donor's weighted entries in order).
  // Stride-1 receivers join it too where the multi-donor receivers outnumber
  // them: there the donor-ordered list reads them better than their own
  // launch (a 200-patch nested batch: 0.326 instead of 0.347 ms per
  // exchange), while aligned copies without multi-donor receivers keep their
  // coalesced launch (a 64-patch jet: 0.0229 instead of 0.0268 ms).
  const auto s1 = single_by_stride.find(1);
  const size_t n_s1 = s1 == single_by_stride.end() ? 0 : s1->second.size();
  const int merge_min = !multi.empty() && multi.size() >= n_s1 ? 1 : 2;
  auto merged = [&](int stride) { return stride >= merge_min; };
  auto upload_lists = [&](drn_ctx::ExchangeLists& x, bool skip_fused) -> int {
    for (auto& kv : single_by_stride) {
      if (merged(kv.first)) continue;
      drn_ctx::ExchangeLists::SingleClass c{};
      c.stride = kv.first;
      std::vector<uint32_t> rec, idx;
      std::vector<double> w;
      for (const auto& e : kv.second) {
        if (skip_fused && fused[e.first]) continue;
        rec.push_back(e.first);
        idx.insert(idx.end(), e.second[0].idx.begin(), e.second[0].idx.end());
        w.insert(w.end(), e.second[0].w.begin(), e.second[0].w.end());
      }
      c.n = (uint32_t)rec.size();
      if (c.n == 0) continue;
      const int rc = upload_vec(ctx, &c.rec, rec) || upload_vec(ctx, &c.idx, idx) || upload_weights(ctx, &c.w, w)
                         ? DRN_E_NOMEM
                         : DRN_OK;
      x.single.push_back(c);  // freed with the lists, also after a failed upload
      if (rc) return fail(ctx, DRN_E_NOMEM, "exchange list upload failed");
    }
    // the multi-donor receivers, sorted by their number of entries in bands
    // of 8 (descending), so that the 64 lanes of a wavefront loop over about
    // as many entries each, and within a band by their first donor cell, so
    // that the lanes read neighbouring donor cells: on a 200-patch nested
    // batch 0.356 ms per exchange, against 0.564 sorted by the count alone,
    // 0.387 with bands of 1 and 0.385 by the donor cell alone;
    // wavefront b's entries start at tbase[b], entry q of lane L at + 64 q + L
    std::vector<std::pair<size_t, const std::pair<uint32_t, std::vector<Group>>*>> order;  // (entries, receiver)
    for (const auto& e : multi) {
      if (skip_fused && fused[e.first]) continue;
      size_t n = 0;
      for (const auto& g : e.second) n += g.idx.size();
      order.push_back({n, &e});
    }
    for (const auto& kv : single_by_stride) {
      if (!merged(kv.first)) continue;
      for (const auto& e : kv.second) {
        if (skip_fused && fused[e.first]) continue;
        order.push_back({e.second[0].idx.size(), &e});
      }
    }
    auto first_donor = [](const std::pair<uint32_t, std::vector<Group>>* e) -> uint32_t {
      return e->second.empty() || e->second[0].idx.empty() ? 0u : e->second[0].idx[0];
    };
    std::stable_sort(order.begin(), order.end(), [&](const auto& a, const auto& b) {
      const size_t ka = (a.first + 7) / 8, kb = (b.first + 7) / 8;
      if (ka != kb) return ka > kb;
      return first_donor(a.second) < first_donor(b.second);
    });
    std::vector<uint32_t> rec, cnt, tbase, ent, et;
    std::vector<double> w;
    const bool turned = !turns.empty();
    for (size_t b = 0; b < order.size(); b += 64) {
      const size_t nb = std::min<size_t>(64, order.size() - b);
      size_t emax = 0;
      for (size_t L = 0; L < nb; ++L) emax = std::max(emax, order[b + L].first);
      if (ent.size() + 64 * emax > 0xffffffffu) return fail(ctx, DRN_E_RANGE, "exchange lists exceed 2^32 entries");
      tbase.push_back((uint32_t)ent.size());
      ent.resize(ent.size() + 64 * emax, 0u);
      w.resize(ent.size(), 0.0);
      if (turned) et.resize(ent.size(), 0u);
      for (size_t L = 0; L < nb; ++L) {
        const auto& e = *order[b + L].second;
        rec.push_back(e.first);
        size_t q = 0;
        for (const auto& g : e.second) {
          for (size_t s2 = 0; s2 < g.idx.size(); ++s2, ++q) {
            const size_t at = tbase.back() + 64 * q + L;
            if (g.idx[s2] & 0x80000000u) return fail(ctx, DRN_E_RANGE, "donor cell index beyond 2^31");
            ent[at] = g.idx[s2] | (s2 + 1 == g.idx.size() ? 0x80000000u : 0u);
            w[at] = g.w[s2];
            if (turned && s2 + 1 == g.idx.size()) et[at] = (uint32_t)g.turn;
          }
        }
        cnt.push_back((uint32_t)q);
      }
    }
    if (!rec.empty()) {
      x.multi_n = (uint32_t)rec.size();
      if (upload_vec(ctx, &x.m_rec, rec) || upload_vec(ctx, &x.m_cnt, cnt) || upload_vec(ctx, &x.m_tbase, tbase) ||
          upload_vec(ctx, &x.m_ent, ent) || upload_weights(ctx, &x.m_w, w))
        return fail(ctx, DRN_E_NOMEM, "exchange list upload failed");
      if (turned && (upload_vec(ctx, &x.m_et, et) || upload_weights(ctx, &x.m_turn, turns)))
        return fail(ctx, DRN_E_NOMEM, "exchange list upload failed");
    }
    return DRN_OK;
  };
  int rc;
  if ((rc = upload_lists(ctx->xfull, false)) || (rc = upload_lists(ctx->xpart, true))) return rc;
  ctx->have_exchange = !ctx->xfull.single.empty() || ctx->xfull.multi_n > 0;
  ctx->n_fused = n_fused;
  HIPCHK(ctx, hipMemcpy(ctx->d_pd, ctx->pd_host.data(), ctx->pd_host.size() * sizeof(PatchDev), hipMemcpyHostToDevice));
  // Dead faces: a face with seek layers whose every boundary cell (the layer the
  // face functor acts on, cartesianiterator.h:190-308) is a receiver of the
  // exchange.  In the RK step the exchange after the stage overwrites those
  // cells, so the functor's flux there never reaches the state.
  {
    std::vector<PatchDev> rk = ctx->pd_host;
    int dead = 0;
    for (const auto& kv : ctx->xt) {
      const int p = kv.first;
      const drn_patch_desc& d = ctx->descs[p];
      std::vector<bool> is_rec((size_t)d.ni * d.nj * d.nk, false);
      for (int64_t r : kv.second.rec_unique) is_rec[(size_t)r] = true;
      const int n3[3] = {d.ni, d.nj, d.nk};
      for (int f = 0; f < 6; ++f) {
        const int a = f >> 1;
        if (d.seek[f] < 1 || n3[a] <= 2) continue;
        const int fixed = (f & 1) ? n3[a] - 1 : 0;
        const int b = (a + 1) % 3, c = (a + 2) % 3;
        bool all = true;
        for (int u = 0; u < n3[b] && all; ++u)
          for (int v = 0; v < n3[c] && all; ++v) {
            int ijk[3];
            ijk[a] = fixed;
            ijk[b] = u;
            ijk[c] = v;
            all = is_rec[((size_t)ijk[0] * d.nj + ijk[1]) * d.nk + ijk[2]];
          }
        if (all) {
          rk[p].bc[f] = BC_DEAD;
          ++dead;
        }
      }
    }
    dfree(ctx->d_pd_rk);
    ctx->d_pd_rk = nullptr;
    ctx->n_dead_faces = dead;
    if (dead > 0) {
      HIPCHK(ctx, dalloc(&ctx->d_pd_rk, rk.size()));
      HIPCHK(ctx, hipMemcpy(ctx->d_pd_rk, rk.data(), rk.size() * sizeof(PatchDev), hipMemcpyHostToDevice));
    }
  }
  return build_chain_list(ctx, push_pairs);
}

int drn_set_exchange_fusion(drn_ctx* ctx, int on) {
  if (!ctx) return DRN_E_ARG;
  if (on != 0 && on != 1) return fail(ctx, DRN_E_ARG, "exchange fusion switch must be 0 or 1");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->fuse = on != 0;
  drop_graph(ctx);
  return DRN_OK;
}

int drn_exchange_fused_cells(const drn_ctx* ctx, int launch, int64_t* n) {
  if (!ctx || !n) return DRN_E_ARG;
  if (launch != DRN_LAUNCH_BATCHED && launch != DRN_LAUNCH_PER_PATCH) return DRN_E_ARG;
  *n = push_active(ctx, launch) ? ctx->n_fused : 0;
  return DRN_OK;
}

int drn_exchange(drn_ctx* ctx, int field) {
  if (!ctx) return DRN_E_ARG;
  int rc;
  if ((rc = check_field(ctx, field))) return rc;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  return do_exchange(ctx, ctx->map[field]);
}

int drn_levelset_set(drn_ctx* ctx, int patch, int64_t n, const int64_t* cells, const double* normals,
                     const int64_t* donors, const double* weights) {
  if (!ctx || n < 0 || (n > 0 && (!cells || !normals || !donors || !weights))) return DRN_E_ARG;
  if (patch < 0 || patch >= (int)ctx->descs.size()) return fail(ctx, DRN_E_ARG, "patch id out of range");
  const int64_t nc = (int64_t)drn_patch_cells(ctx, patch);
  LsTables t;
  t.cells.assign(cells, cells + n);
  t.normals.assign(normals, normals + 3 * n);
  t.donors.assign(donors, donors + 8 * n);
  t.weights.assign(weights, weights + 8 * n);
  for (int64_t a = 0; a < n; ++a)
    if (t.cells[a] < 0 || t.cells[a] >= nc) return fail(ctx, DRN_E_RANGE, "level-set cell outside patch");
  for (int64_t a = 0; a < 8 * n; ++a)
    if (t.donors[a] < 0 || t.donors[a] >= nc) return fail(ctx, DRN_E_RANGE, "mirror donor outside patch");
  ctx->lt[patch] = std::move(t);
  return DRN_OK;
}

int drn_levelset_commit(drn_ctx* ctx) {
  if (!ctx) return DRN_E_ARG;
  if (!ctx->laid_out) return fail(ctx, DRN_E_STATE, "drn_layout not called");
  drop_graph(ctx);  // a captured step holds the old fill lists
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  free_ls(ctx);
  std::vector<uint32_t> cells, donors;
  std::vector<double> nrm, w;
  for (auto& kv : ctx->lt) {
    const uint64_t off = ctx->offs[kv.first];
    const LsTables& t = kv.second;
    for (size_t a = 0; a < t.cells.size(); ++a) cells.push_back((uint32_t)(off + t.cells[a]));
    for (size_t a = 0; a < t.donors.size(); ++a) donors.push_back((uint32_t)(off + t.donors[a]));
    nrm.insert(nrm.end(), t.normals.begin(), t.normals.end());
    w.insert(w.end(), t.weights.begin(), t.weights.end());
  }
  ctx->ls_n = (uint32_t)cells.size();
  if (ctx->ls_n == 0) return DRN_OK;
  int rc;
  if ((rc = upload_vec(ctx, &ctx->ls_cells, cells)) || (rc = upload_vec(ctx, &ctx->ls_donors, donors)) ||
      (rc = upload_weights(ctx, &ctx->ls_nrm, nrm)) || (rc = upload_weights(ctx, &ctx->ls_w, w)))
    return rc;
  HIPCHK(ctx, hipMalloc(&ctx->ls_acc, (size_t)5 * ctx->ls_n * ctx->esize() + 16));
  return DRN_OK;
}

int drn_set_levelset_op(drn_ctx* ctx, int op) {
  if (!ctx) return DRN_E_ARG;
  if (op != DRN_LS_SLIP && op != DRN_LS_SWALL) return fail(ctx, DRN_E_ARG, "unknown level-set operator");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->ls_op = op;
  drop_graph(ctx);  // a captured step holds the fill kernels
  return DRN_OK;
}

int drn_levelset_fill(drn_ctx* ctx, int field, double relax) {
  if (!ctx) return DRN_E_ARG;
  int rc;
  if ((rc = check_field(ctx, field))) return rc;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  return do_ls_fill(ctx, ctx->map[field], relax);
}

int drn_rk_step(drn_ctx* ctx, double dt, int n_alpha, const double* alpha, int launch) {
  if (!ctx || n_alpha < 1 || !alpha) return DRN_E_ARG;
  if (!ctx->laid_out) return fail(ctx, DRN_E_STATE, "drn_layout not called");
  if (launch != DRN_LAUNCH_BATCHED && launch != DRN_LAUNCH_PER_PATCH) return fail(ctx, DRN_E_ARG, "unknown launch form");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  return do_rk_step(ctx, dt, n_alpha, alpha, launch);
}

int drn_rk_graph_run(drn_ctx* ctx, double dt, int n_alpha, const double* alpha, int n_steps) {
  if (!ctx || n_alpha < 1 || !alpha || n_steps < 0) return DRN_E_ARG;
  if (!ctx->laid_out) return fail(ctx, DRN_E_STATE, "drn_layout not called");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (n_steps == 0) return DRN_OK;
  int rc;
  // A replay starts from an exchanged state, so that the captured steps may
  // skip the leading exchange as drn_rk_step does; it runs eagerly when needed.
  if (ctx->exch_buf != ctx->map[0]) {
    if ((rc = do_exchange(ctx, ctx->map[0]))) return rc;
  }
  // A step permutes the three buffers; the steps until the map returns to the
  // start (two for an odd stage count, one for an even one) form the cycle
  // that is captured, once per starting map.
  int map0[NBUF];
  memcpy(map0, ctx->map, sizeof(map0));
  int cyc = 0;
  {
    int m[NBUF];
    memcpy(m, ctx->map, sizeof(m));
    for (cyc = 1; cyc <= 6; ++cyc) {
      const int X = m[0], Y = m[1] == X ? m[2] : m[1];
      int Z = 0;
      while (Z == X || Z == Y) ++Z;
      int cur = X;
      for (int s = 0; s < n_alpha; ++s) cur = (s % 2 == 0) ? Y : Z;
      m[0] = cur;
      m[1] = X;
      m[2] = (cur == Y) ? Z : Y;
      if (memcmp(m, map0, sizeof(m)) == 0) break;
    }
  }
  int done = 0;
  if (n_steps >= cyc) {
    const std::vector<double> al(alpha, alpha + n_alpha);
    int gi = -1;
    for (size_t q = 0; q < ctx->graphs.size(); ++q) {
      const drn_ctx::GraphEntry& e = ctx->graphs[q];
      if (memcmp(e.map, map0, sizeof(map0)) == 0 && e.dt == dt && e.alpha == al && e.steps == cyc) gi = (int)q;
    }
    if (gi < 0) {
      // graphs of another step size or other coefficients are not replayed again
      if (!ctx->graphs.empty() && (ctx->graphs[0].dt != dt || ctx->graphs[0].alpha != al)) drop_graph(ctx);
      const bool was_timing = ctx->timing;
      const int exch0 = ctx->exch_buf;
      ctx->timing = false;
      hipGraph_t graph;
      HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      rc = DRN_OK;
      for (int s = 0; s < cyc && rc == DRN_OK; ++s) rc = do_rk_step(ctx, dt, n_alpha, alpha, DRN_LAUNCH_BATCHED);
      hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
      const int exch_end = ctx->exch_buf;
      ctx->timing = was_timing;
      memcpy(ctx->map, map0, sizeof(map0));
      ctx->exch_buf = exch0;
      if (rc) {
        if (e == hipSuccess) (void)hipGraphDestroy(graph);
        return rc;
      }
      if (e != hipSuccess) return fail(ctx, DRN_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
      drn_ctx::GraphEntry g;
      memcpy(g.map, map0, sizeof(map0));
      g.dt = dt;
      g.alpha = al;
      g.steps = cyc;
      g.exch_end = exch_end;
      g.exec = nullptr;
      e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) return fail(ctx, DRN_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
      ctx->graphs.push_back(g);
      gi = (int)ctx->graphs.size() - 1;
    }
    const drn_ctx::GraphEntry& g = ctx->graphs[(size_t)gi];
    for (; done + cyc <= n_steps; done += cyc) {
      HIPCHK(ctx, hipGraphLaunch(g.exec, ctx->stream));
      ctx->exch_buf = g.exch_end;
    }
  }
  // the remainder runs eagerly and advances the map
  for (; done < n_steps; ++done) {
    if ((rc = do_rk_step(ctx, dt, n_alpha, alpha, DRN_LAUNCH_BATCHED))) return rc;
  }
  return DRN_OK;
}

template <typename T>
static int sums_t(drn_ctx* ctx, const T* U, double* out) {
  const int nt = (int)ctx->tiles.size();
  hipLaunchKernelGGL(k_sums_partial<T>, dim3(nt), dim3(256), 0, ctx->stream, ctx->d_pd, ctx->d_tiles, U, ctx->vstride,
                     ctx->d_part);
  hipLaunchKernelGGL(k_sums_final, dim3(1), dim3(256), 0, ctx->stream, (uint32_t)nt, ctx->d_part, ctx->d_out);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipMemcpyAsync(out, ctx->d_out, 5 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return DRN_OK;
}

int drn_core_sums(drn_ctx* ctx, int field, double* out) {
  if (!ctx || !out) return DRN_E_ARG;
  int rc;
  if ((rc = check_field(ctx, field))) return rc;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const void* U = ctx->buf[ctx->map[field]];
  return ctx->prec == DRN_F32 ? sums_t(ctx, (const float*)U, out) : sums_t(ctx, (const double*)U, out);
}

// drn_core_sums with the partials in `scratch` (through its device address
// ctx->scratch) and the result left in device memory: the same kernels, so the
// same sums bit for bit.  Every partial row read by k_sums_final is written by
// k_sums_partial of the same call, so nothing needs clearing.
template <typename T>
static void sums_noalloc_t(drn_ctx* ctx, const T* U, double* d_out) {
  const int nt = (int)ctx->tiles.size();
  hipLaunchKernelGGL(k_sums_partial<T>, dim3(nt), dim3(256), 0, ctx->stream, ctx->d_pd, ctx->d_tiles, U, ctx->vstride,
                     ctx->scratch);
  hipLaunchKernelGGL(k_sums_final, dim3(1), dim3(256), 0, ctx->stream, (uint32_t)nt, ctx->scratch, d_out);
}

int drn_core_sums_noalloc(drn_ctx* ctx, int field, double* d_out) {
  if (!ctx || !d_out) return DRN_E_ARG;
  int rc;
  if ((rc = check_field(ctx, field))) return rc;
  if (!ctx->scratch) return fail(ctx, DRN_E_STATE, "drn_core_sums_noalloc: no scratch address for this device");
  if (ctx->tiles.empty()) return fail(ctx, DRN_E_STATE, "drn_core_sums_noalloc: no cells");
  if (ctx->tiles.size() > kScratchChunks)
    return fail(ctx, DRN_E_RANGE,
                "drn_core_sums_noalloc: " + std::to_string(ctx->tiles.size()) + " plane chunks exceed the scratch of " +
                    std::to_string(kScratchChunks) + " (use drn_core_sums)");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const void* U = ctx->buf[ctx->map[field]];
  if (ctx->prec == DRN_F32)
    sums_noalloc_t(ctx, (const float*)U, d_out);
  else
    sums_noalloc_t(ctx, (const double*)U, d_out);
  HIPCHK(ctx, hipGetLastError());
  return DRN_OK;
}

// out[0..4] as k_diag_partial's partials, reduced over all tiles
template <typename T>
static int diag_t(drn_ctx* ctx, const T* U, double dt, double* out) {
  const int nt = (int)ctx->tiles.size();
  hipLaunchKernelGGL(k_diag_partial<T>, dim3(nt), dim3(256), 0, ctx->stream, ctx->d_pd, ctx->d_tiles, U, ctx->vstride,
                     dt, ctx->d_part);
  HIPCHK(ctx, hipGetLastError());
  std::vector<double> h((size_t)nt * 5);
  HIPCHK(ctx, hipMemcpyAsync(h.data(), ctx->d_part, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  double o[5] = {0, 1e300, -1e300, 0, 0};
  for (int b = 0; b < nt; ++b) {
    o[0] = std::max(o[0], h[5 * b]);
    o[1] = std::min(o[1], h[5 * b + 1]);
    o[2] = std::max(o[2], h[5 * b + 2]);
    o[3] += h[5 * b + 3];
    o[4] = std::max(o[4], h[5 * b + 4]);
  }
  memcpy(out, o, sizeof(o));
  return DRN_OK;
}

static int diag(drn_ctx* ctx, int field, double dt, double* o5) {
  int rc;
  if ((rc = check_field(ctx, field))) return rc;
  if (!ctx->laid_out) return fail(ctx, DRN_E_STATE, "drn_layout not called");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const void* U = ctx->buf[ctx->map[field]];
  return ctx->prec == DRN_F32 ? diag_t(ctx, (const float*)U, dt, o5) : diag_t(ctx, (const double*)U, dt, o5);
}

int drn_diagnostics(drn_ctx* ctx, int field, double dt, double* out) {
  if (!ctx || !out) return DRN_E_ARG;
  double o[5];
  int rc = diag(ctx, field, dt, o);
  if (rc) return rc;
  memcpy(out, o, 4 * sizeof(double));
  return DRN_OK;
}

int drn_cfl_axes(drn_ctx* ctx, int field, double dt, double* cfl) {
  if (!ctx || !cfl) return DRN_E_ARG;
  double o[5];
  int rc = diag(ctx, field, dt, o);
  if (rc) return rc;
  *cfl = o[4];
  return DRN_OK;
}

template <typename T>
static int vardiff_t(drn_ctx* ctx, int patch, const T* a, const T* b, double* max_norm, double* l2_norm) {
  const int t0 = ctx->tile_first[patch], nt = ctx->tile_first[patch + 1] - t0;
  if (nt > 0)
    hipLaunchKernelGGL(k_vardiff_partial<T>, dim3(nt), dim3(256), 0, ctx->stream, ctx->d_pd, ctx->d_tiles + t0, a, b,
                       ctx->d_part);
  HIPCHK(ctx, hipGetLastError());
  std::vector<double> h((size_t)nt * 2);
  if (nt > 0)
    HIPCHK(ctx, hipMemcpyAsync(h.data(), ctx->d_part, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  double mx = 0, sum = 0;
  for (int q = 0; q < nt; ++q) {
    mx = std::max(mx, h[2 * q]);
    sum += h[2 * q + 1];
  }
  // the maximum is a state-type value: its square root in the state type, as
  // the reference takes it; the sum was accumulated in FP64
  *max_norm = (double)std::sqrt((T)mx);
  *l2_norm = std::sqrt(sum);
  return DRN_OK;
}

int drn_variable_difference(drn_ctx* ctx, int patch, int field1, int var1, int field2, int var2, double* max_norm,
                            double* l2_norm) {
  if (!ctx || !max_norm || !l2_norm) return DRN_E_ARG;
  if (!ctx->laid_out) return fail(ctx, DRN_E_STATE, "drn_layout not called");
  if (patch < 0 || patch >= (int)ctx->descs.size()) return fail(ctx, DRN_E_ARG, "patch id out of range");
  if (var1 < 0 || var1 > 4 || var2 < 0 || var2 > 4) return fail(ctx, DRN_E_ARG, "variable index out of range");
  int rc;
  if ((rc = check_field(ctx, field1)) || (rc = check_field(ctx, field2))) return rc;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t es = ctx->esize();
  const char* A = (const char*)ctx->buf[ctx->map[field1]] + (size_t)var1 * ctx->vstride * es;
  const char* Bv = (const char*)ctx->buf[ctx->map[field2]] + (size_t)var2 * ctx->vstride * es;
  return ctx->prec == DRN_F32 ? vardiff_t(ctx, patch, (const float*)A, (const float*)Bv, max_norm, l2_norm)
                              : vardiff_t(ctx, patch, (const double*)A, (const double*)Bv, max_norm, l2_norm);
}

template <typename T>
static int post_t(drn_ctx* ctx, const T* U, int patch, int mask, double ref_p, double ref_T, void* host) {
  const auto& d = ctx->descs[patch];
  const uint32_t n = (uint32_t)((size_t)d.ni * d.nj * d.nk);
  int nq = 0;
  for (int q = 0; q < 8; ++q) nq += (mask >> q) & 1;
  const size_t bytes = (size_t)nq * n * sizeof(T);
  if (bytes > ctx->post_bytes) {
    dfree(ctx->d_post);
    ctx->d_post = nullptr;
    ctx->post_bytes = 0;
    HIPCHK(ctx, hipMalloc(&ctx->d_post, bytes));
    ctx->post_bytes = bytes;
  }
  hipLaunchKernelGGL(k_post<T>, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, (uint64_t)ctx->offs[patch], U,
                     ctx->vstride, mask, (T)ref_p, (T)ref_T, (T*)ctx->d_post);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipMemcpyAsync(host, ctx->d_post, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return DRN_OK;
}

int drn_post_variables(drn_ctx* ctx, int field, int patch, int mask, double ref_p, double ref_T, void* host_out) {
  if (!ctx || !host_out) return DRN_E_ARG;
  int rc;
  if ((rc = check_field(ctx, field))) return rc;
  if (patch < 0 || patch >= (int)ctx->descs.size()) return fail(ctx, DRN_E_ARG, "patch id out of range");
  if (mask <= 0 || mask > 255) return fail(ctx, DRN_E_ARG, "variable mask must select 1..8 of the 8 quantities");
  if (!(ref_p > 0 && ref_T > 0)) return fail(ctx, DRN_E_ARG, "reference pressure and temperature must be positive");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const void* U = ctx->buf[ctx->map[field]];
  return ctx->prec == DRN_F32 ? post_t(ctx, (const float*)U, patch, mask, ref_p, ref_T, host_out)
                              : post_t(ctx, (const double*)U, patch, mask, ref_p, ref_T, host_out);
}

int drn_set_tiling(drn_ctx* ctx, int launch, int choice) {
  if (!ctx) return DRN_E_ARG;
  if (choice < 0 || choice >= kNumTilings) return fail(ctx, DRN_E_ARG, "unknown tiling choice");
  if (launch != DRN_LAUNCH_BATCHED && launch != DRN_LAUNCH_PER_PATCH) return fail(ctx, DRN_E_ARG, "unknown launch form");
  if (!ctx->laid_out) return fail(ctx, DRN_E_STATE, "drn_layout not called");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  int& slot = launch == DRN_LAUNCH_BATCHED ? ctx->tiling : ctx->tiling_pp;
  const int old = slot;
  slot = choice;
  int rc = build_march(ctx);
  if (rc) {
    slot = old;
    build_march(ctx);
    return rc;
  }
  drop_graph(ctx);
  return DRN_OK;
}

int drn_get_tiling(const drn_ctx* ctx, int launch) {
  if (!ctx) return DRN_E_ARG;
  return launch == DRN_LAUNCH_BATCHED ? ctx->tiling : ctx->tiling_pp;
}

int drn_num_tilings(void) { return kNumTilings; }

// Times every tiling choice on one stage of the given launch form (field 0 ->
// field 2 with field 1 as base, which leaves fields 0 and 1 untouched) and keeps
// the fastest.
int drn_autotune(drn_ctx* ctx, int launch, int* choice) {
  if (!ctx) return DRN_E_ARG;
  if (!ctx->laid_out) return fail(ctx, DRN_E_STATE, "drn_layout not called");
  if (launch != DRN_LAUNCH_BATCHED && launch != DRN_LAUNCH_PER_PATCH) return fail(ctx, DRN_E_ARG, "unknown launch form");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const bool was_timing = ctx->timing;
  ctx->timing = false;
  hipEvent_t e0, e1;
  HIPCHK(ctx, hipEventCreate(&e0));
  HIPCHK(ctx, hipEventCreate(&e1));
  int best = drn_get_tiling(ctx, launch);
  float best_ms = 1e30f;
  int rc = DRN_OK;
  std::vector<std::pair<int, int>> tried;  // (effective kind, list order) already timed as the same
  for (int c = 0; c < kNumTilings && rc == DRN_OK; ++c) {
    const int ek = effective_kind(ctx, c);
    // choices that fall back to the same cell kernel on the same list are one candidate
    if (ek == K_CELLZ || ek == K_CELL || ek == K_XM) {
      const std::pair<int, int> key(ek, kTilings[c].xcd + 4 * kTilings[c].irange);
      bool seen = false;
      for (const auto& t : tried) seen = seen || t == key;
      if (seen) continue;
      tried.push_back(key);
    }
    if (drn_set_tiling(ctx, launch, c) != DRN_OK) continue;
    // a stage and the exchange after it: the cell kernels fill the fused
    // receivers and leave the exchange less to do
    const bool push = push_active(ctx, launch);
    auto trial = [&]() {
      int r = do_stage(ctx, ctx->map[0], ctx->map[1], ctx->map[2], 0.0, launch, push, true);
      return r ? r : do_exchange(ctx, ctx->map[2], push);
    };
    rc = trial();  // warm-up
    if (rc) break;
    HIPCHK(ctx, hipEventRecord(e0, ctx->stream));
    const int reps = 3;
    for (int r = 0; r < reps && rc == DRN_OK; ++r) rc = trial();
    HIPCHK(ctx, hipEventRecord(e1, ctx->stream));
    HIPCHK(ctx, hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(ctx, hipEventElapsedTime(&ms, e0, e1));
    if (ms < best_ms) {
      best_ms = ms;
      best = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  ctx->timing = was_timing;
  if (rc) return rc;
  rc = drn_set_tiling(ctx, launch, best);
  if (choice) *choice = best;
  return rc;
}

int drn_set_arith(drn_ctx* ctx, int mode) {
  if (!ctx) return DRN_E_ARG;
  if (mode != DRN_ARITH_IEEE && mode != DRN_ARITH_FAST && mode != -1) return fail(ctx, DRN_E_ARG, "unknown arithmetic mode");
  ctx->arith = mode;
  if (ctx->laid_out) {
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    int rc = build_march(ctx);
    if (rc) return rc;
  }
  drop_graph(ctx);
  return DRN_OK;
}

int drn_timing_enable(drn_ctx* ctx, int on) {
  if (!ctx) return DRN_E_ARG;
  ctx->timing = on != 0;
  return DRN_OK;
}

int drn_timing(drn_ctx* ctx, int which, double* avg_ms, int64_t* count) {
  if (!ctx || which < 0 || which > 3) return DRN_E_ARG;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  TimerSlot& s = ctx->timers[which];
  double tot = 0;
  int64_t n = 0;
  for (auto& e : s.used) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, e.first, e.second) == hipSuccess) {
      tot += ms;
      ++n;
    }
    s.pool.push_back(e);
  }
  s.used.clear();
  if (avg_ms) *avg_ms = n ? tot / n : 0.0;
  if (count) *count = n;
  return DRN_OK;
}

int drn_sync(drn_ctx* ctx) {
  if (!ctx) return DRN_E_ARG;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return DRN_OK;
}

