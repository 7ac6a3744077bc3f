// Batched ETDRK4 ensembles of 2D vorticity problems on mixed power-of-two
// grids.  No counterpart in the reference (it steps one problem at a time);
// each member is the same computation as the single-grid fused path.
//
// One launch per pass covers every member: the row kernel and each column
// stage kernel read a work list of (member, first row / first column) items
// and run the member's pass with the FFT length templated on its grid size.
// A workgroup of BD threads holds BD/T teams (T = n/P threads per
// transform), so small grids get many rows/columns per workgroup and large
// grids fewer; every team of a workgroup works on the same member, which
// keeps the team barriers uniform.  Items are ordered largest grid first.
//
// Steps are captured once into a HIP graph (10 steps = 80 kernel nodes) and
// replayed.  A variant fixes the workgroup size BD in {128, 256} and, per
// pass (rows, columns), the grid size from which a transform takes 16 points
// per thread instead of 8 (128, 256, 512 or never) -- so each member size
// gets its radix schedule (8^a 2^b ... or 16^a 8^b ...) and tile shape (BD P/n
// rows or columns per workgroup).  The threshold is a template parameter: a
// kernel holds only the bodies it runs, so small members keep the occupancy
// of their 8-point bodies.  The default is the single-grid rule (16 points
// from n = 256 up); the autotuner times all 32 variants on the ensemble
// itself and keeps the fastest.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <vector>

#include "vort2d_ens.h"

namespace ffhip {

constexpr int kEnsChunk = 10;  // steps per captured graph

}  // namespace ffhip

using namespace ffhip;

#ifndef FF_ENS_STREAMS
#define FF_ENS_STREAMS 2  // the step graph runs the members as this many groups on as many streams
#endif

struct EnsVariant {
  int bd = 128;
  int tcol = kEnsDefaultThr, trow = kEnsDefaultThr;  // sizes from these up take 16 points per thread
  bool pair = false;  // one launch per stage pair (tcol == trow, 256 or 2048)
};

struct ffhip_ensemble {
  int prec = FFHIP_F32;
  int count = 0;
  double dt = 0;
  std::vector<VortGeom> geo;
  std::vector<EnsMember> hmem;
  std::vector<void*> Efull;  // exp(dt L) per member, FP64, full rfft layout (export of aliased modes)
  std::map<int, void*> tw;   // twiddles exp(-2 pi i k/n) per length
  void* pool = nullptr;      // all member buffers
  void* dmem = nullptr;      // EnsMember[count]
  int autotune = 0, tuned = 0;
  EnsVariant var;
  double tuned_ms = 0;
  void* items = nullptr;  // int2[ncol + nrow] for var: columns of group A, of group B, rows of A, of B
  int ncol = 0, nrow = 0;
  int ngroup = 1;                          // member groups (the graph runs each on its own stream)
  int col0[FF_ENS_STREAMS + 1] = {}, row0[FF_ENS_STREAMS + 1] = {};  // first column / row item of each group
  size_t lds_col = 0, lds_row = 0;
  hipStream_t cap = nullptr;
  hipStream_t capg[FF_ENS_STREAMS] = {};  // capture streams of groups 1.. (group 0 runs on cap)
  hipEvent_t ev_fork = nullptr, ev_join[FF_ENS_STREAMS] = {};
  hipGraphExec_t exec = nullptr;
  int items_gen = 0, graph_gen = -1;  // the graph holds the item-table pointer it was captured with
  void* sync = nullptr;    // stage-pair counters: rows_done[count], cols_done[count], err
  void* counts = nullptr;  // int2[count]: (row items, column items) per member for var
  double bytes[6] = {0, 0, 0, 0, 0, 0};  // algorithmic bytes per pass over all members
};

namespace ffhip {

static bool pow2i(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }

static bool variant_ok(const ffhip_ensemble* e, EnsVariant v) {
  if (v.pair && (v.tcol != v.trow || (v.tcol != 256 && v.tcol != 2048))) return false;
  for (const auto& g : e->geo)
    if (!ens_ok(g.nx, v.trow, v.bd) || !ens_ok(g.ny, v.tcol, v.bd)) return false;
  return true;
}

// work lists and LDS sizes of a variant
static int build_items(ffhip_ensemble* e, EnsVariant v) {
  std::vector<int> order(e->count);
  for (int m = 0; m < e->count; ++m) order[m] = m;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return (int64_t)e->geo[a].nx * e->geo[a].ny > (int64_t)e->geo[b].nx * e->geo[b].ny;
  });
  std::vector<std::vector<int2>> gcols(FF_ENS_STREAMS), grows(FF_ENS_STREAMS);
  std::vector<int2> cnt(e->count, make_int2(0, 0));
  size_t lc = 0, lr = 0;
  const size_t cs = cplx_size(e->prec);
  for (int im = 0; im < (int)order.size(); ++im) {
    const int m = order[im];
    // members are dealt to the groups in size order (balanced work)
    std::vector<int2>& mc = gcols[im % FF_ENS_STREAMS];
    std::vector<int2>& mr = grows[im % FF_ENS_STREAMS];
    const VortGeom& g = e->geo[m];
    const int Py = ens_pts(g.ny, v.tcol), Px = ens_pts(g.nx, v.trow);
    const int cpb = v.bd / (g.ny / Py), rpb = v.bd / (g.nx / Px);
    FF_CHECK(cpb >= 1 && rpb >= 1, FFHIP_ERR_UNSUPPORTED, "variant %d/%d/%d cannot hold a %d x %d member", v.bd,
             v.tcol, v.trow, g.nx, g.ny);
    for (int c = 0; c < g.TX * kVCB; c += cpb) mc.push_back(make_int2(m, c));
    for (int y = 0; y < g.ny; y += rpb) mr.push_back(make_int2(m, y));
    cnt[m] = make_int2((g.ny + rpb - 1) / rpb, (g.TX * kVCB + cpb - 1) / cpb);
    lc = std::max(lc, ((size_t)cpb * (size_t)team_slice(g.ny + (g.ny >> FF_LDS_PAD_SHIFT), (int)cs, cpb) + tw_slots_rt(g.ny, Py)) * cs);
    lr = std::max(lr, ((size_t)rpb * (size_t)team_slice(g.nx + (g.nx >> FF_LDS_PAD_SHIFT), (int)cs, rpb) + tw_slots_rt(g.nx, Px)) * cs);
  }
  FF_CHECK(lc <= (size_t)max_lds_per_block() && lr <= (size_t)max_lds_per_block(), FFHIP_ERR_UNSUPPORTED,
           "ensemble variant needs more LDS than a workgroup has");
  dev_free(e->items);
  e->items = nullptr;
  std::vector<int2> cols, rows;
  e->ngroup = std::min(FF_ENS_STREAMS, e->count);
  for (int gi = 0; gi < FF_ENS_STREAMS; ++gi) {
    e->col0[gi] = (int)cols.size();
    e->row0[gi] = (int)rows.size();
    cols.insert(cols.end(), gcols[gi].begin(), gcols[gi].end());
    rows.insert(rows.end(), grows[gi].begin(), grows[gi].end());
  }
  e->col0[FF_ENS_STREAMS] = (int)cols.size();
  e->row0[FF_ENS_STREAMS] = (int)rows.size();
  std::vector<int2> all(cols);
  all.insert(all.end(), rows.begin(), rows.end());
  FF_TRY(dev_alloc(&e->items, all.size() * sizeof(int2)));
  FF_HIP(hipMemcpy(e->items, all.data(), all.size() * sizeof(int2), hipMemcpyHostToDevice));
  if (!e->counts) FF_TRY(dev_alloc(&e->counts, (size_t)e->count * sizeof(int2)));
  FF_HIP(hipMemcpy(e->counts, cnt.data(), cnt.size() * sizeof(int2), hipMemcpyHostToDevice));
  e->ncol = (int)cols.size();
  e->nrow = (int)rows.size();
  e->lds_col = lc;
  e->lds_row = lr;
  e->var = v;
  e->items_gen += 1;
  for (int w = 0; w < 6; ++w) e->bytes[w] = 0;
  for (const auto& g : e->geo) {
    for (int st = 0; st <= 4; ++st)
      e->bytes[st] += e->prec == FFHIP_F32 ? col_bytes<float>(g, st, 1) : col_bytes<double>(g, st, 1);
    e->bytes[5] += e->prec == FFHIP_F32 ? row_bytes<float>(g) : row_bytes<double>(g);
  }
  return FFHIP_OK;
}

template <typename K>
static int set_lds(K kern, size_t bytes) {
  FF_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)bytes));
  return FFHIP_OK;
}

// which: 0..4 column stage, 5 row pass.  attr_only: set the kernel's LDS
// limit without launching (done outside any capture).
// part: a group's index, or -1 for every member
template <typename R, int BD, int TC, int TR>
static int ens_pass(ffhip_ensemble* e, int which, hipStream_t st, bool attr_only, int part) {
  const EnsMember* mem = reinterpret_cast<const EnsMember*>(e->dmem);
  const int2* items = reinterpret_cast<const int2*>(e->items);
  static const char* names[6] = {"ens_col_stage0", "ens_col_stage1", "ens_col_stage2",
                                 "ens_col_stage3", "ens_col_stage4", "ens_row"};
  const int c0 = part < 0 ? 0 : e->col0[part], nc = part < 0 ? e->ncol : e->col0[part + 1] - c0;
  const int r0 = part < 0 ? 0 : e->row0[part], nr = part < 0 ? e->nrow : e->row0[part + 1] - r0;
  if (which == 5) {
    auto k = k_ens_row<R, BD, TR>;
    if (attr_only) return set_lds(k, e->lds_row);
    if (nr == 0) return FFHIP_OK;
    prof_launch(names[5], e->bytes[5], k, dim3(nr), dim3(BD), e->lds_row, st, mem, items + e->ncol + r0);
    FF_LAUNCHED("k_ens_row");
    return FFHIP_OK;
  }
  void (*k)(const EnsMember*, const int2*) = nullptr;
  switch (which) {
    case 0: k = k_ens_col<R, BD, TC, 0>; break;
    case 1: k = k_ens_col<R, BD, TC, 1>; break;
    case 2: k = k_ens_col<R, BD, TC, 2>; break;
    case 3: k = k_ens_col<R, BD, TC, 3>; break;
    default: k = k_ens_col<R, BD, TC, 4>; break;
  }
  if (attr_only) return set_lds(k, e->lds_col);
  if (nc == 0) return FFHIP_OK;
  prof_launch(names[which], e->bytes[which], k, dim3(nc), dim3(BD), e->lds_col, st, mem, items + c0);
  FF_LAUNCHED("k_ens_col");
  return FFHIP_OK;
}

template <typename R>
static int ens_pass_v(ffhip_ensemble* e, int which, hipStream_t st, bool attr_only, int part) {
  // column passes use tcol, the row pass trow: instantiate each pass with its own threshold only
  const EnsVariant& v = e->var;
  const int t = which == 5 ? v.trow : v.tcol;
#define FF_ENS_T(BD, T)                                                           \
  (which == 5 ? ens_pass<R, BD, 2048, T>(e, which, st, attr_only, part)         \
              : ens_pass<R, BD, T, 2048>(e, which, st, attr_only, part))
#define FF_ENS_BD(BD)                                                           \
  switch (t) {                                                                  \
    case 128: return FF_ENS_T(BD, 128);                                         \
    case 256: return FF_ENS_T(BD, 256);                                         \
    case 512: return FF_ENS_T(BD, 512);                                         \
    default: return FF_ENS_T(BD, 2048);                                         \
  }
  if (v.bd == 128) {
    FF_ENS_BD(128)
  }
  FF_ENS_BD(256)
#undef FF_ENS_BD
#undef FF_ENS_T
}

static int ens_launch(ffhip_ensemble* e, int which, hipStream_t st, bool attr_only = false, int part = -1) {
  return e->prec == FFHIP_F32 ? ens_pass_v<float>(e, which, st, attr_only, part)
                              : ens_pass_v<double>(e, which, st, attr_only, part);
}

static EnsPairArgs pair_args(const ffhip_ensemble* e, int stage) {
  EnsPairArgs pa;
  pa.mem = reinterpret_cast<const EnsMember*>(e->dmem);
  pa.items = reinterpret_cast<const int2*>(e->items);
  pa.ncol = e->ncol;
  pa.nrow = e->nrow;
  int* sy = reinterpret_cast<int*>(e->sync);
  pa.sy.rows_done = sy;
  pa.sy.cols_done = sy + e->count;
  pa.sy.err = sy + 2 * e->count;
  pa.sy.counts = reinterpret_cast<const int2*>(e->counts);
  pa.lds = std::max(e->lds_row, e->lds_col);
  pa.bytes = e->bytes[5] + e->bytes[stage];
  return pa;
}

static int ens_step(ffhip_ensemble* e, hipStream_t st) {
  for (int stage = 1; stage <= 4; ++stage) {
    if (e->var.pair) {
      FF_TRY(ens_pair_launch(e->prec, e->var.bd, e->var.tcol, stage, pair_args(e, stage), st, false));
    } else {
      FF_TRY(ens_launch(e, 5, st));
      FF_TRY(ens_launch(e, stage, st));
    }
  }
  return FFHIP_OK;
}

// one step of every group, group g on sts[g] (per-pass launches)
static int ens_step_split(ffhip_ensemble* e, const hipStream_t* sts) {
  for (int stage = 1; stage <= 4; ++stage)
    for (int gi = 0; gi < e->ngroup; ++gi) {
      FF_TRY(ens_launch(e, 5, sts[gi], false, gi));
      FF_TRY(ens_launch(e, stage, sts[gi], false, gi));
    }
  return FFHIP_OK;
}

// the stage-pair counters start at zero (a fill kernel, outside any capture)
static __global__ void k_zero_ints(int* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

static int ens_import(ffhip_ensemble* e, void* const* sols, hipStream_t st) {
  for (int m = 0; m < e->count; ++m) FF_TRY(vort_import(e->prec, e->geo[m], sols[m], e->hmem[m].ca.U, st));
  return FFHIP_OK;
}

static void drop_graph(ffhip_ensemble* e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  e->exec = nullptr;
}

static int capture_graph(ffhip_ensemble* e) {
  drop_graph(e);
  if (!e->cap) FF_HIP(hipStreamCreateWithFlags(&e->cap, hipStreamNonBlocking));
  // member groups on their own streams (forked from and joined back into
  // the capture stream): one group's passes fill the others' launch gaps and
  // partly filled tails
  const bool split = !e->var.pair && e->ngroup > 1;
  hipStream_t sts[FF_ENS_STREAMS] = {};
  if (split) {
    sts[0] = e->cap;
    for (int gi = 1; gi < e->ngroup; ++gi) {
      if (!e->capg[gi]) FF_HIP(hipStreamCreateWithFlags(&e->capg[gi], hipStreamNonBlocking));
      if (!e->ev_join[gi]) FF_HIP(hipEventCreateWithFlags(&e->ev_join[gi], hipEventDisableTiming));
      sts[gi] = e->capg[gi];
    }
    if (!e->ev_fork) FF_HIP(hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming));
  }
  FF_HIP(hipStreamBeginCapture(e->cap, hipStreamCaptureModeThreadLocal));
  int rc = FFHIP_OK;
  if (split) {
    hipError_t ef = hipEventRecord(e->ev_fork, e->cap);
    for (int gi = 1; gi < e->ngroup && ef == hipSuccess; ++gi) ef = hipStreamWaitEvent(sts[gi], e->ev_fork, 0);
    if (ef != hipSuccess) rc = fail(FFHIP_ERR_HIP, "ensemble graph fork: %s", hipGetErrorString(ef));
    for (int k = 0; k < kEnsChunk && rc == FFHIP_OK; ++k) rc = ens_step_split(e, sts);
    for (int gi = 1; gi < e->ngroup; ++gi) {  // join every group back (also after an error: the capture must end)
      hipError_t ej = hipEventRecord(e->ev_join[gi], sts[gi]);
      if (ej == hipSuccess) ej = hipStreamWaitEvent(e->cap, e->ev_join[gi], 0);
      if (ej != hipSuccess && rc == FFHIP_OK) rc = fail(FFHIP_ERR_HIP, "ensemble graph join: %s", hipGetErrorString(ej));
    }
  } else {
    for (int k = 0; k < kEnsChunk && rc == FFHIP_OK; ++k) rc = ens_step(e, e->cap);
  }
  hipGraph_t graph = nullptr;
  hipError_t ee = hipStreamEndCapture(e->cap, &graph);
  if (rc != FFHIP_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (ee != hipSuccess) return fail(FFHIP_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(ee));
  ee = hipGraphInstantiate(&e->exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ee != hipSuccess) {
    e->exec = nullptr;
    return fail(FFHIP_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(ee));
  }
  e->graph_gen = e->items_gen;
  FF_HIP(hipGraphUpload(e->exec, e->cap));  // the first replay then carries no upload
  return FFHIP_OK;
}

// LDS limits of every pass of the current variant
static int set_variant_attrs(ffhip_ensemble* e) {
  for (int w = 0; w <= 5; ++w) FF_TRY(ens_launch(e, w, nullptr, true));
  if (e->var.pair)
    for (int stage = 1; stage <= 4; ++stage)
      FF_TRY(ens_pair_launch(e->prec, e->var.bd, e->var.tcol, stage, pair_args(e, stage), nullptr, true));
  return FFHIP_OK;
}

// ms per step of variant v on the ensemble's own buffers (their contents are
// overwritten by the import that follows), or a negative value if v is invalid
static int time_variant(ffhip_ensemble* e, EnsVariant v, hipStream_t st, hipEvent_t a, hipEvent_t b, double* ms) {
  *ms = -1;
  if (!variant_ok(e, v)) return FFHIP_OK;
  FF_HIP(hipStreamSynchronize(st));  // the item table is replaced below
  FF_TRY(build_items(e, v));
  FF_TRY(set_variant_attrs(e));
  int* sync = reinterpret_cast<int*>(e->sync);
  if (v.pair) {
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(256), 0, st, sync, 2 * e->count + 1);
    FF_LAUNCHED("k_zero_ints");
  }
  for (int k = 0; k < 2; ++k) FF_TRY(ens_step(e, st));
  FF_HIP(hipEventRecord(a, st));
  for (int k = 0; k < 5; ++k) FF_TRY(ens_step(e, st));
  FF_HIP(hipEventRecord(b, st));
  FF_HIP(hipEventSynchronize(b));
  float t = 0;
  FF_HIP(hipEventElapsedTime(&t, a, b));
  *ms = t / 5.0;
  if (v.pair) {  // a variant whose waits ran out is not a candidate
    int err = 0;
    FF_HIP(hipMemcpy(&err, sync + 2 * e->count, sizeof(int), hipMemcpyDeviceToHost));
    if (err) *ms = -1;
  }
  static const bool log = getenv("FFHIP_ENS_TUNE_LOG") && atoi(getenv("FFHIP_ENS_TUNE_LOG")) != 0;
  if (log)
    fprintf(stderr, "ensemble variant bd=%d tcol=%d trow=%d pair=%d: %.4f ms/step\n", v.bd, v.tcol, v.trow,
            (int)v.pair, *ms);
  return FFHIP_OK;
}

// Every (BD, column threshold, row threshold) that fits the members, with one
// launch per pass, and every (BD, threshold) with one launch per stage pair;
// the fastest is kept.
static int autotune_search(ffhip_ensemble* e, hipStream_t st, hipEvent_t a, hipEvent_t b) {
  double best = 1e300;
  EnsVariant bestv = e->var;
  for (int bd : {128, 256})
    for (int tc : kEnsThresholds)
      for (int tr : kEnsThresholds) {
        EnsVariant v;
        v.bd = bd;
        v.tcol = tc;
        v.trow = tr;
        double ms = 0;
        FF_TRY(time_variant(e, v, st, a, b, &ms));
        if (ms >= 0 && ms < best) {
          best = ms;
          bestv = v;
        }
      }
  for (int bd : {128, 256})  // one launch per stage pair
    for (int t : {256, 2048}) {
      EnsVariant v;
      v.bd = bd;
      v.tcol = v.trow = t;
      v.pair = true;
      double ms = 0;
      FF_TRY(time_variant(e, v, st, a, b, &ms));
      if (ms >= 0 && ms < best) {
        best = ms;
        bestv = v;
      }
    }
  FF_CHECK(best < 1e300, FFHIP_ERR_UNSUPPORTED, "autotune: no variant fits the ensemble");
  FF_HIP(hipStreamSynchronize(st));
  FF_TRY(build_items(e, bestv));
  e->tuned_ms = best;
  e->tuned = 1;
  return FFHIP_OK;
}

static int autotune(ffhip_ensemble* e, hipStream_t st) {
  hipEvent_t a, b;
  FF_HIP(hipEventCreate(&a));
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return fail(FFHIP_ERR_HIP, "autotune: hipEventCreate failed");
  }
  const int rc = autotune_search(e, st, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return rc;
}

static void ens_free(ffhip_ensemble* e) {
  if (!e) return;
  drop_graph(e);
  if (e->cap) (void)hipStreamDestroy(e->cap);
  for (int gi = 0; gi < FF_ENS_STREAMS; ++gi) {
    if (e->capg[gi]) (void)hipStreamDestroy(e->capg[gi]);
    if (e->ev_join[gi]) (void)hipEventDestroy(e->ev_join[gi]);
  }
  if (e->ev_fork) (void)hipEventDestroy(e->ev_fork);
  dev_free(e->pool);
  dev_free(e->dmem);
  dev_free(e->items);
  dev_free(e->sync);
  dev_free(e->counts);
  for (void* p : e->Efull) dev_free(p);
  for (auto& kv : e->tw) dev_free(kv.second);
  delete e;
}

}  // namespace ffhip

extern "C" {

int ffhip_ensemble_create(ffhip_ensemble** out, int prec, int count, const int64_t* n, const double* box,
                          double aliased_fraction, const double* const* L, double dt, int autotune_flag) {
  clear_error();
  FF_CHECK(out, FFHIP_ERR_ARG, "null output pointer");
  *out = nullptr;
  FF_CHECK(valid_prec(prec), FFHIP_ERR_ARG, "bad precision %d", prec);
  FF_CHECK(count >= 1 && n && box && L, FFHIP_ERR_ARG, "need count >= 1 and non-null n, box, L");
  FF_CHECK(aliased_fraction > 0 && aliased_fraction < 1, FFHIP_ERR_ARG,
           "ensembles need 0 < aliased_fraction < 1 (the fused passes keep the dealiased band only)");
  ffhip_ensemble* e = new (std::nothrow) ffhip_ensemble();
  FF_CHECK(e, FFHIP_ERR_OOM, "host allocation failed");
  e->prec = prec;
  e->count = count;
  e->dt = prec == FFHIP_F32 ? (double)(float)dt : dt;
  e->autotune = autotune_flag;
  int rc = FFHIP_OK;
  const size_t cs = cplx_size(prec);
  size_t total = 0;
  std::vector<size_t> off(count);
  const double twopi = 6.283185307179586476925286766559;
  for (int m = 0; m < count && rc == FFHIP_OK; ++m) {
    const int64_t nx = n[2 * m], ny = n[2 * m + 1];
    if (!pow2i(nx) || !pow2i(ny) || nx < 64 || ny < 64 || nx > 1024 || ny > 1024) {
      rc = fail(FFHIP_ERR_UNSUPPORTED, "member %d: ensemble grids must be powers of two in 64..1024 (got %lld x %lld)",
                m, (long long)nx, (long long)ny);
      break;
    }
    if (!(box[2 * m] > 0 && box[2 * m + 1] > 0)) {
      rc = fail(FFHIP_ERR_ARG, "member %d: domain lengths must be positive", m);
      break;
    }
    int64_t klo, khi, krlo, krhi, llo, lhi;
    rc = ffhip_aliased_wavenumbers(nx, nx / 2 + 1, aliased_fraction, nullptr, nullptr, &krlo, &krhi);
    if (rc == FFHIP_OK) rc = ffhip_aliased_wavenumbers(ny, ny / 2 + 1, aliased_fraction, &llo, &lhi, &klo, &khi);
    if (rc != FFHIP_OK) break;
    VortGeom g;
    g.nx = (int)nx;
    g.ny = (int)ny;
    g.nkr = (int)(nx / 2 + 1);
    g.KX = (int)krlo - 1;
    g.KXp = (g.KX + kVCB - 1) / kVCB * kVCB;
    g.TX = g.KXp / kVCB;
    g.ylo = (int)llo - 1;
    g.yhi = (int)lhi - 1;
    g.KY = g.ny - (g.yhi - g.ylo + 1);
    g.dk = twopi / box[2 * m];
    g.dl = twopi / box[2 * m + 1];
    g.s2 = 1.0 / ((double)nx * (double)ny * (double)nx * (double)ny);
    if (g.KX < 1 || g.KX > g.nx / 2 || g.ylo < 1 || g.yhi < g.ylo || g.ylo > g.ny / 2 || g.yhi < g.ny / 2) {
      rc = fail(FFHIP_ERR_UNSUPPORTED, "member %d: aliased_fraction leaves no proper kept band", m);
      break;
    }
    e->geo.push_back(g);
    off[m] = total;
    const size_t sb = (size_t)g.TX * g.KY * kVCB * cs, ib = (size_t)g.TX * g.ny * kVCB * cs;
    const size_t cb = (size_t)g.TX * g.KY * kVCB * sizeof(double);
    total += 4 * sb + 5 * ib + 6 * cb;
    total = (total + 255) / 256 * 256;
  }
  if (rc == FFHIP_OK) rc = dev_alloc(&e->pool, total);
  e->hmem.resize(e->geo.size());
  e->Efull.assign(count, nullptr);
  for (int m = 0; m < count && rc == FFHIP_OK; ++m) {
    const VortGeom& g = e->geo[m];
    const size_t sb = (size_t)g.TX * g.KY * kVCB * cs, ib = (size_t)g.TX * g.ny * kVCB * cs;
    const size_t nc = (size_t)g.TX * g.KY * kVCB;
    char* p = reinterpret_cast<char*>(e->pool) + off[m];
    EnsMember& em = e->hmem[m];
    em.g = g;
    VortColArgs& ca = em.ca;
    ca.U = p;
    ca.N1 = p + sb;
    ca.N2 = p + 2 * sb;
    ca.A = p + 3 * sb;
    char* q = p + 4 * sb;
    ca.I0 = q;
    ca.I1 = q + ib;
    ca.I2 = q + 2 * ib;
    ca.O0 = q + 3 * ib;
    ca.O1 = q + 4 * ib;
    double* pc = reinterpret_cast<double*>(q + 5 * ib);
    ca.Z = pc;
    ca.Al = pc + nc;
    ca.Be = pc + 2 * nc;
    ca.Ga = pc + 3 * nc;
    ca.E = pc + 4 * nc;
    ca.Eh = pc + 5 * nc;
    ca.do_inverse = 1;
    ca.Fl = nullptr;
    for (int len : {g.nx, g.ny}) {
      if (!e->tw.count(len)) {
        void* t = nullptr;
        rc = upload_roots(len, len, prec, &t);
        if (rc != FFHIP_OK) break;
        e->tw[len] = t;
      }
    }
    if (rc != FFHIP_OK) break;
    ca.tw = e->tw[g.ny];
    VortRowArgs& ra = em.ra;
    ra.I0 = ca.I0;
    ra.I1 = ca.I1;
    ra.I2 = ca.I2;
    ra.O0 = q + 3 * ib;
    ra.O1 = q + 4 * ib;
    ra.tw = e->tw[g.nx];
    ra.F = nullptr;
    // coefficients: the single-grid stepper's own construction, then packed
    ffhip_etdrk4* ts = nullptr;
    const int64_t dims[2] = {g.nkr, g.ny};
    if (!L[m]) {
      rc = fail(FFHIP_ERR_ARG, "member %d: null L", m);
      break;
    }
    rc = ffhip_etdrk4_create(&ts, L[m], nullptr, 2, dims, dt, prec);
    if (rc != FFHIP_OK) break;
    const int64_t ncnt = (int64_t)nc;
    for (int c = 0; c < 6; ++c) {
      hipLaunchKernelGGL(k_vpack, dim3(blocks_for(ncnt, 256, 4096)), dim3(256), 0, 0,
                         reinterpret_cast<const double*>(ts->coef[c]), pc + c * nc, g);
    }
    // keep exp(dt L) in the full layout for the aliased modes at export
    rc = dev_alloc(&e->Efull[m], (size_t)g.nkr * g.ny * sizeof(double));
    if (rc == FFHIP_OK) {
      hipError_t he = hipMemcpy(e->Efull[m], ts->coef[4], (size_t)g.nkr * g.ny * sizeof(double),
                                hipMemcpyDeviceToDevice);
      if (he != hipSuccess) rc = fail(FFHIP_ERR_HIP, "hipMemcpy failed: %s", hipGetErrorString(he));
    }
    if (rc == FFHIP_OK) {
      hipError_t he = hipDeviceSynchronize();
      if (he != hipSuccess) rc = fail(FFHIP_ERR_HIP, "coefficient packing failed: %s", hipGetErrorString(he));
    }
    ffhip_etdrk4_destroy(ts);
  }
  if (rc == FFHIP_OK) rc = dev_alloc(&e->sync, (2 * (size_t)count + 1) * sizeof(int));  // zero-filled
  if (rc == FFHIP_OK) rc = dev_alloc(&e->dmem, e->hmem.size() * sizeof(EnsMember));
  if (rc == FFHIP_OK) {
    hipError_t he = hipMemcpy(e->dmem, e->hmem.data(), e->hmem.size() * sizeof(EnsMember), hipMemcpyHostToDevice);
    if (he != hipSuccess) rc = fail(FFHIP_ERR_HIP, "hipMemcpy failed: %s", hipGetErrorString(he));
  }
  if (rc == FFHIP_OK) {
    EnsVariant v;  // the single-grid kernels' radix schedule
    if (!variant_ok(e, v)) v.bd = 256;
    // FFHIP_ENS_VARIANT="bd,tcol,trow[,pair]" (thresholds 128, 256, 512 or 2048; pair 0 or 1)
    // fixes the untuned variant
    if (const char* ev = getenv("FFHIP_ENS_VARIANT")) {
      EnsVariant f;
      int pair = 0;
      auto thr_ok = [](int t) { return t == 128 || t == 256 || t == 512 || t == 2048; };
      const int nf = sscanf(ev, "%d,%d,%d,%d", &f.bd, &f.tcol, &f.trow, &pair);
      f.pair = pair != 0;
      if (nf < 3 || (f.bd != 128 && f.bd != 256) || !thr_ok(f.tcol) || !thr_ok(f.trow) || !variant_ok(e, f))
        rc = fail(FFHIP_ERR_ARG, "FFHIP_ENS_VARIANT=%s is not a valid variant for this ensemble", ev);
      v = f;
    }
    if (rc == FFHIP_OK) rc = build_items(e, v);
  }
  if (rc != FFHIP_OK) {
    ens_free(e);
    return rc;
  }
  *out = e;
  return FFHIP_OK;
}

int ffhip_ensemble_destroy(ffhip_ensemble* e) {
  clear_error();
  ens_free(e);
  return FFHIP_OK;
}

int ffhip_ensemble_run(ffhip_ensemble* e, void* const* sols, int64_t nsteps, void* stream) {
  clear_error();
  FF_CHECK(e && sols, FFHIP_ERR_ARG, "null argument");
  FF_CHECK(nsteps >= 0, FFHIP_ERR_ARG, "nsteps must be >= 0");
  for (int m = 0; m < e->count; ++m) FF_CHECK(sols[m], FFHIP_ERR_ARG, "member %d: null state", m);
  if (nsteps == 0) return FFHIP_OK;
  hipStream_t st = as_stream(stream);
  if (e->autotune && !e->tuned) FF_TRY(autotune(e, st));
  FF_TRY(set_variant_attrs(e));
  if (e->var.pair) {
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(256), 0, st, reinterpret_cast<int*>(e->sync), 2 * e->count + 1);
    FF_LAUNCHED("k_zero_ints");
  }
  FF_TRY(ens_import(e, sols, st));
  FF_TRY(ens_launch(e, 0, st));
  int64_t left = nsteps;
  const bool graphs = !g_prof_on && nsteps >= kEnsChunk;
  if (graphs) {
    if (!e->exec || e->graph_gen != e->items_gen) FF_TRY(capture_graph(e));
    for (; left >= kEnsChunk; left -= kEnsChunk) FF_HIP(hipGraphLaunch(e->exec, st));
  }
  for (; left > 0; --left) FF_TRY(ens_step(e, st));
  for (int m = 0; m < e->count; ++m)
    FF_TRY(vort_export(e->prec, e->geo[m], sols[m], e->hmem[m].ca.U, reinterpret_cast<const double*>(e->Efull[m]),
                       nullptr, nsteps, st));
  if (e->var.pair) {  // did every column workgroup's wait end normally?
    int err = 0;
    FF_HIP(hipMemcpyAsync(&err, reinterpret_cast<int*>(e->sync) + 2 * e->count, sizeof(int), hipMemcpyDeviceToHost,
                          st));
    FF_HIP(hipStreamSynchronize(st));
    FF_CHECK(err == 0, FFHIP_ERR_HIP, "ensemble stage-pair launch: a column workgroup's wait for its rows ran out");
  }
  return FFHIP_OK;
}

int ffhip_ensemble_info(const ffhip_ensemble* e, int* threads, int* prule, double* ms_per_step) {
  clear_error();
  FF_CHECK(e, FFHIP_ERR_ARG, "null ensemble");
  if (threads) *threads = e->var.bd;
  if (prule) *prule = ens_mask(e->var.tcol) | (ens_mask(e->var.trow) << 8) | (e->var.pair ? 1 << 16 : 0);
  if (ms_per_step) *ms_per_step = e->tuned ? e->tuned_ms : 0.0;
  return FFHIP_OK;
}

}  // extern "C"
