// Host side of the MI355X HSC solver: workspaces, setup pipeline, persistent
// PCG launch (single and batched), autotuner, acceptance checks and the C ABI
// declared in include/hsc.h.  Mirrors LinearHSC::do_build / setup / do_solve
// of the reference (reproduction/hsc_backend/linear_hsc.cu:172-436).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/hsc.h"
#include "common.hpp"
#include "kernels.hpp"
#include "learned.hpp"
#include "mlp.hpp"

#include <atomic>
#include <cstring>
#include <fstream>
#include <functional>
#include <memory>
#include <sstream>

#include <sys/stat.h>

namespace {
std::atomic<long long> g_alloc_count{0};
}  // namespace

void hsc::note_alloc() { g_alloc_count.fetch_add(1, std::memory_order_relaxed); }

namespace {

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

struct Buf {
  void* p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes == 0) bytes = 64;
    if (bytes <= cap) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    hipError_t e = hipMalloc(&p, bytes);
    if (e == hipSuccess) {
      cap = bytes;
      hsc::note_alloc();
    }
    return e;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
  ~Buf() {
    if (p) (void)hipFree(p);
  }
};

// Pinned host staging: device-to-host copies into it are asynchronous, so
// several results are fetched with one stream synchronisation.
struct HostBuf {
  void* p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
    if (e == hipSuccess) {
      cap = bytes;
      hsc::note_alloc();
    }
    return e;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
  ~HostBuf() {
    if (p) (void)hipHostFree(p);
  }
};

// Scratch of one persistent PCG launch: system descriptors and outputs,
// chunk tables, reduction partials, grid-barrier counters, error flag.
struct PcgScratch {
  Buf sysdesc, sysout, chunks, partA, partE, partF, partL, partR, barrier, errflag, flags, ztag;
  int nflags = 0;  // row-tile pass flags zeroed before each launch (0: point-to-point pass sync off)
  size_t ztag_bytes = 0;  // pass pairs' tagged values zeroed before each launch (0: no pairs)
  HostBuf host;    // pinned copies of the outputs and the error flag
  // last uploaded descriptor and chunk tables, pinned: the uploads read
  // them asynchronously, so they are rewritten only once up_ev has passed
  // (an identical plan is not uploaded again)
  HostBuf up_sys, up_ch;
  size_t up_sys_n = 0, up_ch_n = 0;
  void* up_sys_ptr = nullptr;
  void* up_ch_ptr = nullptr;
  hipEvent_t up_ev = nullptr;
  bool up_pending = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  ~PcgScratch() {
    if (up_pending) (void)hipEventSynchronize(up_ev);
    if (up_ev) (void)hipEventDestroy(up_ev);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
  }
};

struct HscError {
  int code;
  std::string msg;
};

#define HIPCHK(expr)                                                                      \
  do {                                                                                    \
    hipError_t _e = (expr);                                                               \
    if (_e != hipSuccess)                                                                 \
      throw HscError{HSC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)};     \
  } while (0)

constexpr long long kTimeoutTicks = 100LL * 1000 * 1000 * 20;  // 20 s at 100 MHz

struct DeviceInfo {
  int cu = 0;
  int lds_max = 65536;  // LDS bytes one workgroup may allocate
};

DeviceInfo device_info(int dev) {
  static std::mutex mu;
  static std::map<int, DeviceInfo> cache;
  std::lock_guard<std::mutex> g(mu);
  auto it = cache.find(dev);
  if (it != cache.end()) return it->second;
  hipDeviceProp_t prop;
  DeviceInfo d;
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess) {
    d.cu = prop.multiProcessorCount;
    if (prop.sharedMemPerBlock > 0) d.lds_max = (int)prop.sharedMemPerBlock;
  }
  if (d.cu <= 0) d.cu = 1;
  cache[dev] = d;
  return d;
}

// autotune table: size class -> launch shape.  The class is the device (CU
// count), log2 of the nodes, blocks per row, rank / 16, precision, batch or
// single, preconditioner family (Neumann; Chebyshev and learned share the
// polynomial passes), preconditioner order and pass sync.
struct TuneKey {
  int cu, nb_log2, deg, kbucket, mixed, batch, poly, order, psync;
  auto tie() const { return std::tie(cu, nb_log2, deg, kbucket, mixed, batch, poly, order, psync); }
  bool operator<(const TuneKey& o) const { return tie() < o.tie(); }
};
std::mutex g_tune_mu;
struct TuneVal {
  int G, wgpcu, sched;
};
std::map<TuneKey, TuneVal> g_tune;
// the file the table is loaded from and appended to ("" = memory only)
std::string g_tune_path;
bool g_tune_path_set = false, g_tune_loaded = false;
constexpr const char* kTuneMagic = "hsc-tune-v4";

std::string default_tune_path() {
  if (const char* e = std::getenv("HSC_TUNE_FILE")) return e;
  if (const char* home = std::getenv("HOME")) return std::string(home) + "/.cache/hsc_mi355x/tune_v4.txt";
  return "";
}

// Load the table file once (caller holds g_tune_mu).  Lines:
// "hsc-tune-v4 cu nb_log2 deg kbucket mixed batch poly order psync G wgpcu sched";
// anything else is skipped.
void load_tune_locked() {
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  if (!g_tune_path_set) {
    g_tune_path = default_tune_path();
    g_tune_path_set = true;
  }
  if (g_tune_path.empty()) return;
  std::ifstream f(g_tune_path);
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::string magic;
    TuneKey k{};
    TuneVal v{};
    if (!(is >> magic) || magic != kTuneMagic) continue;
    if (!(is >> k.cu >> k.nb_log2 >> k.deg >> k.kbucket >> k.mixed >> k.batch >> k.poly >> k.order >> k.psync >> v.G >>
          v.wgpcu >> v.sched))
      continue;
    if (!(v.G == 4 || v.G == 8 || v.G == 16) || v.wgpcu < 1 || v.wgpcu > 4 || v.sched < 1 || v.sched > 2) continue;
    g_tune[k] = v;
  }
}

// Append one entry to the table file (caller holds g_tune_mu); an unwritable
// path leaves the table in memory only.
void save_tune_locked(const TuneKey& k, const TuneVal& v) {
  if (g_tune_path.empty()) return;
  // create the missing parent directories (mkdir -p of the path's directory)
  for (size_t i = g_tune_path.find('/', 1); i != std::string::npos; i = g_tune_path.find('/', i + 1))
    (void)::mkdir(g_tune_path.substr(0, i).c_str(), 0755);
  FILE* f = std::fopen(g_tune_path.c_str(), "a");
  if (!f) return;
  std::fprintf(f, "%s %d %d %d %d %d %d %d %d %d %d %d %d\n", kTuneMagic, k.cu, k.nb_log2, k.deg, k.kbucket, k.mixed,
               k.batch, k.poly, k.order, k.psync, v.G, v.wgpcu, v.sched);
  std::fclose(f);
}

}  // namespace

struct hsc_handle {
  hsc_config cfg{};
  int device = 0;
  std::string err;
  // system layout
  int N = 0, qo = 0, qm = 0, fo = 0, fn = 0, nb = 0, nnzb = 0, a = 0, k = 0, limit = 0;
  long long count = 0;
  bool blockdiag = true, setup_ok = false;
  double aca_err = 0.0, abd_res = 0.0, setup_ms = 0.0;
  int doublings = 0;
  int abd_mode = 0;                 // ABD inverse that ran: 0 exact, 1 Neumann polynomial, 2 with doubling
  unsigned long long gen = 0;       // bumped by every setup and reconfiguration (captured batches check it)
  bool launch_ready = false;        // G/wgpcu/sched chosen and the launch scratch sized for them
  int G = 8, wgpcu = 1, sched = 1;
  bool pass_sync_used = false;  // the last PCG launch synchronised its passes per tile
  bool pipe_used = false;       // the last PCG launch ran the pipelined loop
  bool pairs_used = false;      // the last PCG launch ran pass pairs (set up for them; after a solve, as run)
  int max_coupled128 = 0;  // most coupled nodes in one 128-node tile
  // row-tile neighbour lists for 128, 64 and 32 rows per tile (lanes per row 4, 8, 16)
  Buf tnbr[4], tnbr_bits;  // [3]: 4 lanes per row over the body-aligned tiles
  bool tnbr_ok[4] = {false, false, false, false};
  // body-aligned row tiles of the latency schedule at 4 lanes per row
  // (ntiles 0: none, the fixed 128-row tiles apply) and the most coupled
  // nodes in one of them
  Buf tstart, reach;
  int ntiles = 0, max_coupled_tiles = 0;
  // pass pairs (128-row tiles): halo packs, built for pack capacity
  // pairs_hbl (-1: not built since the last setup), over the tile table or not
  Buf halo;
  int pairs_hbl = -1;
  bool pairs_tab = false;
  // chunk tables of the streaming operator kernels (strip_nch -1: not built
  // since the last setup)
  Buf hbits, hrow;
  int strip_nch = -1;
  double lambda_max = 0.0;  // power-iteration estimate of lambda_max(Dinv D) (Chebyshev weights)
  double omegas[hsc::kMaxPolySteps] = {};  // per-pass preconditioner weights
  double* xsol = nullptr;  // FEM part of the caller's x during a solve
  double* last_xsol = nullptr;  // that of the last solve (only its address is kept)
  // buffers
  Buf row_count, rowptr, cursor, active_flag, act_prefix, act_node, act_dof;
  Buf key_col, key_src, colidx, val64, val32, invD64, invD32;
  Buf A, E, R0, R1, U, V, hist, pivots, acapart, acacand, Binv, P, s1, s2, s3, W1, C, Lm, W64, W32;
  Buf acab, acabi;  // the ACA-by-body path's scratch (doubles, ints)
  Buf rhs, r, p0, p1, z0, z1, q, q64, yfull, small, ybuf, misc, ints;
  // persistent-launch scratch
  Buf barrier, errflag, respart, profbuf;  // barrier/errflag: setup kernels (ACA)
  PcgScratch pcg;                           // the persistent PCG launch of hsc_solve / hsc_solve_batched
  // learned preconditioner: graph MLP over D's node graph
  Buf learned, nbr_ptr, nbr, dfeat, sdiag, xfeat, omlp, lpred;
  double poly_ratio = 0.0;       // interval ratio b / a of the polynomial passes (Chebyshev or learned; else 0)
  double learned_out[3] = {0.0, 0.0, 0.0};  // the network's channel means at the last setup
  int learned_used = 0;          // the last setup's passes follow the network's interval
  int fallback_reason = 0;       // 2: the network's output was rejected (Neumann weights used)
  Buf pwr;  // power iteration (Chebyshev weights)
  long long learned_bytes = 0;
  long long learned_version = 0;       // bumped by each hsc_set_learned_weights
  long long mlp_weights_version = -1;  // the version the network's bf16 copies were cast from
  std::unique_ptr<hsc::MlpNet> mlp;
  // setup: a second stream runs the D-side work (row sort, value gather,
  // power iteration) beside ACA; pinned staging for the setup's small results
  hipStream_t aux = nullptr;
  hipEvent_t aux_ev = nullptr;
  // learned: a third stream runs the network on D beside the power iteration
  hipStream_t aux2 = nullptr;
  hipEvent_t d_ev = nullptr;
  // the end of each D-side chain, waited for by the main stream, and the
  // end of D's column sort there
  hipEvent_t aux_done = nullptr, aux2_done = nullptr, sort_ev = nullptr;
  HostBuf stage;
  Buf stage_dev;  // the setup's small results gathered on the device, in the pinned copy's layout
  ~hsc_handle() {
    if (aux) (void)hipStreamSynchronize(aux);
    if (aux2) (void)hipStreamSynchronize(aux2);
    if (aux_ev) (void)hipEventDestroy(aux_ev);
    if (d_ev) (void)hipEventDestroy(d_ev);
    if (aux_done) (void)hipEventDestroy(aux_done);
    if (aux2_done) (void)hipEventDestroy(aux2_done);
    if (sort_ev) (void)hipEventDestroy(sort_ev);
    if (aux) (void)hipStreamDestroy(aux);
    if (aux2) (void)hipStreamDestroy(aux2);
  }
};

namespace {

void set_err(hsc_handle* h, int code, const std::string& m) {
  if (h) h->err = m;
  (void)code;
}

int validate_config(const hsc_config& c, std::string& why) {
  // linear_hsc.cu:188-189
  if (c.max_rank > hsc::kMaxRankLds) {
    why = "max_rank must be at most 256 (coupling rank held in LDS by the Schur PCG)";
    return HSC_ERR_INVALID_ARGUMENT;
  }
  if (!(c.max_rank > 0 && c.neumann_order > 0 && c.max_iter > 0 && c.omega > 0 && c.omega <= 1)) {
    why = "Invalid HSC parameters (max_rank, neumann_order, max_iter > 0 and 0 < omega <= 1 required)";
    return HSC_ERR_INVALID_ARGUMENT;
  }
  if (!(c.tol_rate > 0) || !(c.abd_inverse_tol > 0) || !(c.full_residual_limit > 0) || c.aca_rtol < 0) {
    why = "Invalid HSC tolerances";
    return HSC_ERR_INVALID_ARGUMENT;
  }
  if (c.abd_inverse < 0 || c.abd_inverse > 2 || c.precision < 0 || c.precision > 1 || c.preconditioner < 0 ||
      c.preconditioner > 2 || c.residual_replace_every < 0) {
    why = "Invalid HSC mode selection";
    return HSC_ERR_INVALID_ARGUMENT;
  }
  if (!(c.lanes_per_row == 0 || c.lanes_per_row == 4 || c.lanes_per_row == 8 || c.lanes_per_row == 16)) {
    why = "lanes_per_row must be 0, 4, 8 or 16";
    return HSC_ERR_INVALID_ARGUMENT;
  }
  if (c.wg_per_cu < 0 || c.wg_per_cu > 4) {
    why = "wg_per_cu must be in 0..4";
    return HSC_ERR_INVALID_ARGUMENT;
  }
  if (c.preconditioner == HSC_PRECOND_LEARNED && c.neumann_order > hsc::kMaxPolySteps) {
    why = "learned preconditioner: neumann_order <= 32 required";
    return HSC_ERR_INVALID_ARGUMENT;
  }
  if (c.preconditioner == HSC_PRECOND_CHEBYSHEV &&
      !(c.neumann_order <= hsc::kMaxPolySteps && c.poly_interval_ratio > 1.0)) {
    why = "Chebyshev preconditioner: neumann_order <= 32 and poly_interval_ratio > 1 required";
    return HSC_ERR_INVALID_ARGUMENT;
  }
  if (c.pass_sync < 0 || c.pass_sync > 2) {
    why = "pass_sync must be 0, 1 or 2";
    return HSC_ERR_INVALID_ARGUMENT;
  }
  if (c.pipelined < 0 || c.pipelined > 1) {
    why = "pipelined must be 0 or 1";
    return HSC_ERR_INVALID_ARGUMENT;
  }
  if (c.operator_mapping < 0 || c.operator_mapping > 2) {
    why = "operator_mapping must be 0, 1 or 2";
    return HSC_ERR_INVALID_ARGUMENT;
  }
  if (c.schedule < 0 || c.schedule > 3) {
    why = "schedule must be 0, 1, 2 or 3";
    return HSC_ERR_INVALID_ARGUMENT;
  }
  return HSC_OK;
}

hsc::SysDesc make_desc(hsc_handle* h) {
  hsc::SysDesc S{};
  S.nb = h->nb;
  S.n = h->fn;
  S.a = h->k > 0 ? h->a : 0;
  S.k = h->k;
  S.nnzb = h->nnzb;
  S.rowptr = h->rowptr.as<int>();
  S.colidx = h->colidx.as<int>();
  S.val64 = h->val64.as<double>();
  S.val32 = h->val32.as<float>();
  S.invD64 = h->invD64.as<double>();
  S.invD32 = h->invD32.as<float>();
  S.act_node = h->act_node.as<int>();
  S.act_dof = h->act_dof.as<int>();
  S.W64 = h->W64.as<double>();
  S.W32 = h->W32.as<float>();
  S.rhs = h->rhs.as<double>();
  S.x = h->xsol;
  S.r = h->r.as<double>();
  S.p[0] = h->p0.p;
  S.p[1] = h->p1.p;
  S.z[0] = h->z0.p;
  S.z[1] = h->z1.p;
  S.q = h->q.p;
  S.q64 = h->cfg.precision == HSC_PRECISION_MIXED ? h->q64.as<double>() : h->q.as<double>();
  S.tol = h->cfg.tol_rate;
  S.max_iter = h->cfg.max_iter;
  return S;
}

// chunk tables of the streaming operator kernels (built by the setup)
hsc::StripDesc strip_desc(const hsc_handle* h) {
  const int nch = (h->nnzb + 63) / 64;
  if (h->strip_nch != nch) throw HscError{HSC_ERR_NOT_SETUP, "chunk tables not built"};
  return hsc::StripDesc{h->hbits.as<unsigned long long>(), h->hrow.as<int>(), nch, 0};
}

// Algorithmic bytes of one standalone operator launch (0 FP64 D x, 1 FP64
// pass, 2 FP32 pass): values and column indices, row pointers, x gathered
// once, y written, and for a pass r and the inverse diagonal blocks.
long long operator_bytes(int kernel, long long nb, long long nnzb) {
  const long long n = 3 * nb;
  if (kernel == 0) return 76 * nnzb + 4 * (nb + 1) + 16 * n;
  if (kernel == 1) return 76 * nnzb + 4 * (nb + 1) + 24 * n + 72 * nb;
  return 40 * nnzb + 4 * (nb + 1) + 16 * n + 36 * nb;
}

// Whether a standalone operator runs the streaming kernels.  By working set
// (operator_mapping 0): from 256 MiB on, beyond the 256 MB Infinity Cache,
// the streaming kernels are faster (on 8.3M blocks: 128 vs 171 us for D x);
// below it the row mapping's repeated reads hit the cache and its shorter
// chain of dependent loads wins (on 0.26M blocks: 5.8 vs 9.6 us).
bool streaming_ops(const hsc_handle* h, int kernel) {
  if (h->cfg.operator_mapping == 1) return false;
  if (h->cfg.operator_mapping == 2) return true;
  return operator_bytes(kernel, h->nb, h->nnzb) >= (256ll << 20);
}

// One standalone operator launch with the handle's mapping: row mapping at
// 16 lanes per row (8 for the FP32 pass, measured fastest at each size)
void launch_operator(const hsc_handle* h, int kernel, const hsc::SysDesc& S, const void* x, void* y, double w,
                     hipStream_t st) {
  if (streaming_ops(h, kernel)) HIPCHK(hsc::launch_strip_op(kernel, S, strip_desc(h), x, y, w, st));
  else HIPCHK(hsc::launch_row_op(kernel, kernel == 2 ? 8 : 16, S, x, y, w, st));
}

// XCD-aware chunk order: workgroup b runs on XCD b % 8; chunks are numbered
// so that each XCD owns a contiguous run of chunks (neighbouring rows share
// that XCD's L2).
void xcd_ranks(int b0, int n, std::vector<int>& out) {
  int cnt[8] = {0};
  for (int b = b0; b < b0 + n; ++b) cnt[b % 8]++;
  int base[8];
  int acc = 0;
  for (int x = 0; x < 8; ++x) {
    base[x] = acc;
    acc += cnt[x];
  }
  int seen[8] = {0};
  for (int b = b0; b < b0 + n; ++b) {
    const int x = b % 8;
    out[b] = base[x] + seen[x]++;
  }
}

int tile_of_chunk(int c, int nC, int T) { return (int)(((long long)c * T) / nC); }
int chunk_of(int t, int nC, int T) {
  int c = (int)(((long long)t * nC) / T);
  c = std::min(std::max(c, 0), nC - 1);
  while (c + 1 < nC && tile_of_chunk(c + 1, nC, T) <= t) ++c;
  while (c > 0 && tile_of_chunk(c, nC, T) > t) --c;
  return c;
}

// Launch shape of the persistent PCG: lanes per block row, workgroups per CU,
// and the work split (1 = latency: the grid is sized by the SpMV row tiles,
// the low-rank workgroups are split off by byte share, 128-node elementwise
// tiles; 2 = throughput: the grid covers row and low-rank tiles, split by
// tile count, 512-node elementwise tiles).
struct LaunchCfg {
  int G, wgpcu, sched;
};

// Launch plan of the persistent PCG over a batch of set-up handles: tile and
// chunk tables (uploaded to h0's scratch on `st`), kernel parameters.  The
// host copies of the tables live in the plan until the uploads complete.
struct PcgPlan {
  hsc::PcgParams P{};
  size_t shmem = 0;
  int G = 8;
  bool mixed = false;
  bool res = true;  // latency-schedule kernel
  bool pipe = false;  // the pipelined loop
  std::vector<hsc::SysDesc> S;
  std::vector<int> ch;
};

PcgPlan plan_pcg(hsc_handle* const* hs, int nsys, LaunchCfg lc, int max_iter_override, hipStream_t st,
                 PcgScratch& sc) {
  hsc_handle* h0 = hs[0];
  PcgPlan plan;
  const bool mixed = h0->cfg.precision == HSC_PRECISION_MIXED;
  const int G = lc.G;
  const int R = hsc::kBlock / G;
  const int node_tile = lc.sched == 2 ? hsc::kNodeTileMax : 128;
  plan.G = G;
  plan.mixed = mixed;
  plan.res = lc.sched != 2;
  std::vector<hsc::SysDesc>& S = plan.S;
  S.resize(nsys);
  // body-aligned row tiles: schedule 3 (the latency schedule) at 4 lanes per
  // row (row tiles and 128-node tiles then coincide), standard loop
  const bool tab_ok = lc.sched == 3 && G == 4 && !h0->cfg.pipelined;
  int rt = 0, et = 0, lt = 0, t_total = 0;
  double lr_bytes = 0, row_bytes = 0;
  const double eltM = mixed ? 4.0 : 8.0;
  for (int s = 0; s < nsys; ++s) {
    S[s] = make_desc(hs[s]);
    if (!S[s].x) throw HscError{HSC_ERR_INVALID_ARGUMENT, "solution vector not bound"};
    if (max_iter_override > 0) S[s].max_iter = max_iter_override;
    S[s].tstart = tab_ok && hs[s]->ntiles > 0 ? hs[s]->tstart.as<int>() : nullptr;
    S[s].rt_begin = rt;
    rt += S[s].tstart ? hs[s]->ntiles : (S[s].nb + R - 1) / R;
    S[s].rt_end = rt;
    S[s].et_begin = et;
    et += S[s].tstart ? hs[s]->ntiles : (S[s].nb + node_tile - 1) / node_tile;
    S[s].et_end = et;
    S[s].lt_begin = lt;
    if (S[s].k > 0) lt += (S[s].a + hsc::kLrTile - 1) / hsc::kLrTile;
    S[s].lt_end = lt;
    S[s].t_off = t_total;
    t_total += S[s].k;
    lr_bytes += (double)S[s].a * S[s].k * eltM;
    row_bytes += (double)S[s].nnzb * (9 * eltM + 4);
    if (S[s].k > hsc::kMaxRankLds) throw HscError{HSC_ERR_CAPACITY, "rank exceeds 256"};
  }
  const DeviceInfo di = device_info(h0->device);
  // LDS cache of the coupled W rows of a resident node tile (latency
  // schedule, 128-node tiles): sized by the most coupled nodes in any tile
  int wc_rows = 0;
  size_t wc_bytes = 0;
  const size_t t_bytes = ((size_t)std::max(t_total, 1) * sizeof(double) + 15) / 16 * 16;
  if (lc.sched != 2) {
    int kmax = 0;
    for (int s = 0; s < nsys; ++s) {
      wc_rows = std::max(wc_rows, S[s].tstart ? hs[s]->max_coupled_tiles : hs[s]->max_coupled128);
      kmax = std::max(kmax, hs[s]->k);
    }
    wc_bytes = ((size_t)wc_rows * 3 * kmax * (mixed ? sizeof(float) : sizeof(double)) + 15) / 16 * 16;
    if (wc_rows == 0 || kmax == 0 || wc_bytes > 64 * 1024) {
      wc_rows = 0;
      wc_bytes = 0;
    }
  }
  // dynamic LDS starts with the system descriptors and the node tile's
  // low-rank sums
  const size_t sys_bytes = (sizeof(hsc::SysDesc) * nsys + 15) / 16 * 16;
  const size_t sg_bytes = ((size_t)node_tile * 3 * sizeof(double) + 15) / 16 * 16;
  // the pipelined loop (requested by the config) runs when every row tile
  // and its 128 nodes get their own workgroup: 4 lanes per row, latency
  // split, polynomial passes, recursive-residual stopping
  const bool pipe_req = h0->cfg.pipelined && G == 4 && lc.sched != 2 && node_tile == R && h0->cfg.neumann_order >= 2 &&
                        !h0->cfg.true_residual_every_iteration && hsc::kPipeRows == R;
  size_t shmem = sys_bytes + sg_bytes + t_bytes + wc_bytes;
  int occ = 0;
  HIPCHK(hsc::pcg_occupancy(G, mixed, lc.sched != 2, false, shmem, &occ));
  // (the pipelined loop keeps two more t vectors in LDS: W^T z and W^T x)
  const size_t shmem_pipe = sys_bytes + sg_bytes + 3 * t_bytes + wc_bytes;
  int occ_pipe = 0;
  if (pipe_req) HIPCHK(hsc::pcg_occupancy(G, mixed, true, true, shmem_pipe, &occ_pipe));
  if (pipe_req) occ = std::min(occ, occ_pipe);
  if (occ <= 0) throw HscError{HSC_ERR_CAPACITY, "persistent PCG kernel does not fit on a CU"};
  const int cap = occ * di.cu;
  // Stage A runs row tiles (SpMV) on nA workgroups and low-rank tiles (W^T p)
  // on nL others.  Latency split: the grid is sized by the row tiles and the
  // low-rank share follows the bytes.  Throughput split: both kinds of tile
  // cost about one dependent-load chain (measured on MI355X), so the share
  // follows the tile counts and the grid covers both kinds when the
  // co-resident capacity allows.
  // Aligned latency split: when every row tile gets its own workgroup and
  // room is left, the row workgroups own the same row tile in stage A and in
  // the Neumann passes (so its matrix rows stay in registers), and the
  // remaining workgroups take the low-rank tiles.
  const bool thr = lc.sched == 2;
  const int capw = std::min(lc.wgpcu, occ) * di.cu;
  bool aligned = false;
  int nwg = std::min(capw, std::max(thr ? rt + lt : rt, 1));
  int nL = 0;
  if (!thr && rt + (lt > 0 ? 1 : 0) <= capw) {
    aligned = true;
    nL = lt > 0 ? std::min(lt, capw - rt) : 0;
    nwg = rt + nL;
  } else if (lt > 0) {
    nwg = std::max(nwg, 2);
    nL = thr ? (int)std::lround((double)nwg * lt / (double)(rt + lt))
             : (int)std::lround(nwg * lr_bytes / std::max(lr_bytes + row_bytes, 1.0));
    nL = std::max(1, std::min(nL, std::min(lt, nwg - 1)));
  }
  if (nwg > cap) throw HscError{HSC_ERR_CAPACITY, "grid exceeds co-resident capacity"};
  const int nA = nwg - nL, nF = aligned ? nA : nwg;
  std::vector<int> chA(nwg, -1), chF(nwg, -1), chL(nwg, -1), tmp(nwg, -1);
  if (nL > 0) {
    xcd_ranks(0, nL, tmp);
    for (int b = 0; b < nL; ++b) chL[b] = tmp[b];
  }
  xcd_ranks(nL, nA, tmp);
  for (int b = nL; b < nwg; ++b) chA[b] = tmp[b];
  if (aligned) {
    for (int b = nL; b < nwg; ++b) chF[b] = chA[b];  // low-rank workgroups: no row or node tiles
  } else {
    xcd_ranks(0, nwg, chF);
  }
  int lrp = 0;
  for (int s = 0; s < nsys; ++s) {
    S[s].cA0 = chunk_of(S[s].rt_begin, nA, rt);
    S[s].cA1 = chunk_of(S[s].rt_end - 1, nA, rt);
    S[s].cF0 = chunk_of(S[s].rt_begin, nF, rt);
    S[s].cF1 = chunk_of(S[s].rt_end - 1, nF, rt);
    S[s].cE0 = chunk_of(S[s].et_begin, nF, et);
    S[s].cE1 = chunk_of(S[s].et_end - 1, nF, et);
    if (S[s].k > 0) {
      S[s].cL0 = chunk_of(S[s].lt_begin, nL, lt);
      S[s].cL1 = chunk_of(S[s].lt_end - 1, nL, lt);
      S[s].lrp_off = lrp;
      lrp += (S[s].cL1 - S[s].cL0 + 1) * S[s].k;
    } else {
      S[s].cL0 = 0;
      S[s].cL1 = -1;
      S[s].lrp_off = lrp;
    }
  }
  // point-to-point sync between preconditioner passes (aligned latency
  // split only: every row workgroup owns one resident row tile)
  bool psync = h0->cfg.pass_sync && aligned && !thr && h0->cfg.neumann_order > 2;
  for (int s = 0; s < nsys && psync; ++s) psync = S[s].rt_end - S[s].rt_begin <= hsc::kNbrMaxTiles;
  if (psync) {
    for (int s = 0; s < nsys; ++s) {
      hsc_handle* h = hs[s];
      const int li = S[s].tstart ? 3 : (G == 4 ? 0 : (G == 8 ? 1 : 2));
      if (!h->tnbr_ok[li]) {
        const int rtl = S[s].rt_end - S[s].rt_begin, words = (rtl + 31) / 32;
        HIPCHK(h->tnbr[li].ensure(sizeof(int) * (size_t)rtl * hsc::kNbrStride));
        HIPCHK(h->tnbr_bits.ensure(sizeof(unsigned) * (size_t)rtl * words));
        HIPCHK(hsc::launch_tile_neighbors(h->rowptr.as<int>(), h->colidx.as<int>(), h->nb, R, S[s].tstart, rtl,
                                          h->tnbr_bits.as<unsigned>(), h->tnbr[li].as<int>(), st));
        h->tnbr_ok[li] = true;
      }
      S[s].nbr = h->tnbr[li].as<int>();
    }
    HIPCHK(sc.flags.ensure(sizeof(unsigned) * (size_t)rt));
  }
  const bool pipe = pipe_req && aligned && !thr && nA == rt;
  sc.nflags = psync ? rt : 0;
  sc.ztag_bytes = 0;
  plan.pipe = pipe;
  if (pipe) shmem = shmem_pipe;
  // pass pairs (pass_sync = 2): one system, FP32 resident rows at 4 lanes per
  // row, at least two passes.  The halo region takes the LDS the kernel
  // leaves free, and the pack capacity (hbl blocks) follows from it.
  bool pairs = false;
  int hbl = 0;
  const size_t halo_off = t_bytes + wc_bytes;  // relative to the t vectors (16-byte aligned)
  if (psync && !pipe && h0->cfg.pass_sync == 2 && nsys == 1 && G == 4 && mixed && R == hsc::kHaloTileRows &&
      h0->cfg.neumann_order >= 3) {
    size_t stat = 0;
    HIPCHK(hsc::pcg_static_lds(G, mixed, true, false, &stat));
    const long long avail = (long long)di.lds_max - (long long)stat - (long long)(sys_bytes + sg_bytes + halo_off) -
                            hsc::kHaloScratchBytes - hsc::kHaloOffBlk - 512;
    hbl = (int)std::min<long long>(avail / 38, (long long)4 * hsc::kHaloU * hsc::kHaloRows);
    while (hbl > 0 && (long long)(sys_bytes + sg_bytes + halo_off) + hsc::halo_pack_bytes(hbl) + hsc::kHaloScratchBytes +
                              (long long)stat > (long long)di.lds_max)
      --hbl;
    if (hbl >= 64) {
      const size_t shmem_pairs =
          sys_bytes + sg_bytes + halo_off + (size_t)hsc::halo_pack_bytes(hbl) + hsc::kHaloScratchBytes;
      int occ_pairs = 0;
      HIPCHK(hsc::pcg_occupancy(G, mixed, true, false, shmem_pairs, &occ_pairs));
      if (occ_pairs >= 1 && nwg <= std::min(lc.wgpcu, occ_pairs) * di.cu) {
        pairs = true;
        shmem = shmem_pairs;
      }
    }
  }
  if (pairs) {
    hsc_handle* h = h0;
    const bool tab = S[0].tstart != nullptr;
    if (h->pairs_hbl != hbl || h->pairs_tab != tab) {
      const int rtl = S[0].rt_end - S[0].rt_begin;
      HIPCHK(h->halo.ensure((size_t)hsc::kHaloBaseBytes + (size_t)rtl * (size_t)hsc::halo_pack_bytes(hbl)));
      HIPCHK(hsc::launch_tile_pairs(h->rowptr.as<int>(), h->colidx.as<int>(), h->val32.as<float>(),
                                    h->invD32.as<float>(), h->nb, S[0].tstart, rtl, hbl, h->halo.as<unsigned char>(),
                                    st));
      h->pairs_hbl = hbl;
      h->pairs_tab = tab;
    }
    // two buffers of tagged pass values, zeroed before each launch
    sc.ztag_bytes = sizeof(unsigned long long) * 2 * 3 * (size_t)h->nb;
    HIPCHK(sc.ztag.ensure(sc.ztag_bytes));
    S[0].halo = h->halo.as<unsigned char>();
    S[0].zt = sc.ztag.as<unsigned long long>();
  }
  for (int s = 0; s < nsys; ++s) {
    hs[s]->pass_sync_used = psync;
    hs[s]->pipe_used = pipe;
    hs[s]->pairs_used = pairs;
  }
  HIPCHK(sc.sysdesc.ensure(sizeof(hsc::SysDesc) * nsys));
  HIPCHK(sc.sysout.ensure(sizeof(hsc::SysOut) * nsys));
  HIPCHK(sc.chunks.ensure(sizeof(int) * 3 * nwg));
  HIPCHK(sc.partA.ensure(sizeof(double) * 2 * (size_t)nsys * nA));  // (second half: the pipelined loop's z.q_old)
  HIPCHK(sc.partE.ensure(sizeof(double) * 2 * (size_t)nsys * nF));
  HIPCHK(sc.partF.ensure(sizeof(double) * (size_t)nsys * nF));
  HIPCHK(sc.partL.ensure(sizeof(double) * (size_t)std::max(lrp, 1)));
  if (pipe) HIPCHK(sc.partR.ensure(sizeof(double) * (size_t)nF * std::max(t_total, 1)));
  HIPCHK(sc.barrier.ensure(sizeof(unsigned) * hsc::kBarrierWords));
  HIPCHK(sc.errflag.ensure(64));
  std::vector<int>& ch = plan.ch;
  ch.resize(3 * nwg);
  std::copy(chA.begin(), chA.end(), ch.begin());
  std::copy(chF.begin(), chF.end(), ch.begin() + nwg);
  std::copy(chL.begin(), chL.end(), ch.begin() + 2 * nwg);
  // upload the tables unless the same ones are already in place (a solve
  // of the same handles with the same launch shape reuses them)
  // from pinned staging copies, so the caller need not wait for them
  bool uploaded = false;
  auto upload = [&](void* dst, const void* src, size_t bytes, HostBuf& last, size_t& last_n, void*& last_ptr) {
    if (last_ptr == dst && last_n == bytes && std::memcmp(last.p, src, bytes) == 0) return;
    if (sc.up_pending) {  // an earlier upload may still read the staging copies
      HIPCHK(hipEventSynchronize(sc.up_ev));
      sc.up_pending = false;
    }
    last_ptr = nullptr;
    HIPCHK(last.ensure(bytes));
    std::memcpy(last.p, src, bytes);
    HIPCHK(hipMemcpyAsync(dst, last.p, bytes, hipMemcpyHostToDevice, st));
    last_n = bytes;
    last_ptr = dst;
    uploaded = true;
  };
  upload(sc.sysdesc.p, S.data(), sizeof(hsc::SysDesc) * nsys, sc.up_sys, sc.up_sys_n, sc.up_sys_ptr);
  upload(sc.chunks.p, ch.data(), sizeof(int) * 3 * nwg, sc.up_ch, sc.up_ch_n, sc.up_ch_ptr);
  if (uploaded) {
    if (!sc.up_ev) HIPCHK(hipEventCreateWithFlags(&sc.up_ev, hipEventDisableTiming));
    HIPCHK(hipEventRecord(sc.up_ev, st));
    sc.up_pending = true;
  }
  hsc::PcgParams& P = plan.P;
  P.pass_sync = psync ? 1 : 0;
  P.pass_flags = psync ? sc.flags.as<unsigned>() : nullptr;
  P.sys_bytes = (int)sys_bytes;
  P.sg_bytes = (int)sg_bytes;
  P.halo = pairs ? 1 : 0;
  P.halo_off = (int)halo_off;
  P.hbl = hbl;
  P.halo_stride = pairs ? hsc::halo_pack_bytes(hbl) : 0;
  P.sys = sc.sysdesc.as<hsc::SysDesc>();
  P.out = sc.sysout.as<hsc::SysOut>();
  P.nsys = nsys;
  P.order = h0->cfg.neumann_order;
  {
    // per-pass weights; a Chebyshev batch uses the interval of its largest
    // spectrum bound (an upper bound for every system)
    const hsc_handle* hw = h0;
    for (int s = 1; s < nsys; ++s)
      if (hs[s]->lambda_max > hw->lambda_max) hw = hs[s];
    for (int j = 0; j < hsc::kMaxPolySteps; ++j) P.omegas[j] = hw->omegas[j];
  }
  P.mode_true_every = h0->cfg.true_residual_every_iteration;
  P.replace_every = mixed ? h0->cfg.residual_replace_every : 0;
  P.nwg = nwg;
  P.n_lr_wg = nL;
  P.nA = nA;
  P.nF = nF;
  P.nL = std::max(nL, 1);
  P.rt_total = rt;
  P.et_total = et;
  P.lt_total = lt;
  P.rows_per_tile = R;
  P.node_tile = node_tile;
  P.wc_rows = wc_rows;
  P.wc_off = (int)(pipe ? 3 * t_bytes : t_bytes);
  P.chunkA = sc.chunks.as<int>();
  P.chunkF = sc.chunks.as<int>() + nwg;
  P.chunkL = sc.chunks.as<int>() + 2 * nwg;
  P.partA = sc.partA.as<double>();
  P.partE = sc.partE.as<double>();
  P.partF = sc.partF.as<double>();
  P.partL = sc.partL.as<double>();
  P.partR = pipe ? sc.partR.as<double>() : nullptr;
  P.barrier = sc.barrier.as<unsigned>();
  P.error = sc.errflag.as<int>();
  P.timeout_ticks = kTimeoutTicks;
  P.t_total = t_total;
  P.prof = nullptr;
  plan.shmem = shmem;
  return plan;
}

// Zero the grid-barrier counters, the error flag and every system's p_old
// (multiplied by beta = 0 in the first iteration, so it must be finite).
// Fill kernels, not memsets, so that the sequence can be captured in a graph.
void reset_pcg(hsc_handle* const* hs, int nsys, PcgScratch& sc, hipStream_t st) {
  // one launch per kZeroListMax buffers
  std::vector<const void*> zp = {sc.barrier.p, sc.errflag.p};
  std::vector<size_t> zb = {sizeof(unsigned) * hsc::kBarrierWords, 64};
  if (sc.nflags > 0) {
    zp.push_back(sc.flags.p);
    zb.push_back(sizeof(unsigned) * sc.nflags);
  }
  if (sc.ztag_bytes > 0) {
    zp.push_back(sc.ztag.p);
    zb.push_back(sc.ztag_bytes);
  }
  for (int s = 0; s < nsys; ++s) {
    zp.push_back(hs[s]->p0.p);
    zb.push_back(sizeof(double) * std::max(hs[s]->fn, 1));
  }
  for (size_t o = 0; o < zp.size(); o += hsc::kZeroListMax) {
    const int cnt = (int)std::min<size_t>(hsc::kZeroListMax, zp.size() - o);
    HIPCHK(hsc::launch_zero_multi(zp.data() + o, zb.data() + o, cnt, st));
  }
}

// Pinned host staging of collect_pcg for nsys systems, and the launch's
// timing events.
size_t host_out_offset(int nsys) { return (sizeof(hsc::SysOut) * nsys + 64 + 15) / 16 * 16; }
void ensure_host_scratch(PcgScratch& sc, int nsys) {
  HIPCHK(sc.host.ensure(host_out_offset(nsys) + sizeof(double) * 3 * (size_t)nsys));
  if (!sc.ev0) HIPCHK(hipEventCreate(&sc.ev0));
  if (!sc.ev1) HIPCHK(hipEventCreate(&sc.ev1));
}

// Copy the per-system outputs back (pinned) and wait; raises on a device
// timeout.  With fin_hs, also each handle's acceptance sums (|b - Hx|^2, |b|^2,
// b.x) into *fin, fetched in the same synchronisation.
std::vector<hsc::SysOut> collect_pcg(PcgScratch& sc, int nsys, hipStream_t st, hsc_handle* const* fin_hs = nullptr,
                                     std::vector<double>* fin = nullptr) {
  const size_t out_bytes = sizeof(hsc::SysOut) * nsys;
  const size_t fin_off = host_out_offset(nsys);
  ensure_host_scratch(sc, nsys);
  unsigned char* hb = sc.host.as<unsigned char>();
  int* derr = reinterpret_cast<int*>(hb + out_bytes);
  double* hfin = reinterpret_cast<double*>(hb + fin_off);
  *derr = 0;
  hipError_t ce = hipMemcpyAsync(hb, sc.sysout.p, out_bytes, hipMemcpyDeviceToHost, st);
  if (ce == hipSuccess) ce = hipMemcpyAsync(derr, sc.errflag.p, sizeof(int), hipMemcpyDeviceToHost, st);
  if (fin_hs)
    for (int i = 0; i < nsys && ce == hipSuccess; ++i)
      ce = hipMemcpyAsync(hfin + 3 * i, fin_hs[i]->misc.as<double>() + 20, 3 * sizeof(double),
                          hipMemcpyDeviceToHost, st);
  if (ce == hipSuccess) ce = hipStreamSynchronize(st);
  if (ce != hipSuccess) throw HscError{HSC_ERR_HIP, std::string("PCG: ") + hipGetErrorString(ce)};
  if (*derr) throw HscError{HSC_ERR_DEVICE_TIMEOUT, "persistent PCG: bounded grid wait expired"};
  std::vector<hsc::SysOut> out(nsys);
  std::memcpy(out.data(), hb, out_bytes);
  if (fin_hs && fin) fin->assign(hfin, hfin + 3 * (size_t)nsys);
  if (fin_hs)
    for (int i = 0; i < nsys; ++i) fin_hs[i]->pairs_used = out[i].pairs != 0;
  return out;
}

// Persistent PCG over a batch of set-up handles (one cooperative launch).
// `after` enqueues further work behind the launch before the one host
// synchronisation (the back-substitution and acceptance sums of a solve,
// whose results come back in *fin).
std::vector<hsc::SysOut> run_pcg(hsc_handle* const* hs, int nsys, LaunchCfg lc, int max_iter_override,
                                 hipStream_t st, float* pcg_ms, const std::function<void()>& after = {},
                                 std::vector<double>* fin = nullptr) {
  hsc_handle* h0 = hs[0];
  PcgPlan plan = plan_pcg(hs, nsys, lc, max_iter_override, st, h0->pcg);
  reset_pcg(hs, nsys, h0->pcg, st);
  hsc::PcgParams& P = plan.P;
  const int nwg = P.nwg;
  static const bool stage_timing = std::getenv("HSC_STAGE_TIMING") != nullptr;
  if (stage_timing) {
    HIPCHK(h0->profbuf.ensure(sizeof(long long) * 32 * nwg));
    HIPCHK(hsc::launch_zero(h0->profbuf.p, sizeof(long long) * 32 * nwg, st));
    P.prof = h0->profbuf.as<long long>();
  }
  PcgScratch& sc = h0->pcg;
  ensure_host_scratch(sc, nsys);
  hipEvent_t e0 = sc.ev0, e1 = sc.ev1;
  HIPCHK(hipEventRecord(e0, st));
  hipError_t le = hsc::launch_pcg(plan.G, plan.mixed, plan.res, plan.pipe, P, plan.shmem, st);
  HIPCHK(hipEventRecord(e1, st));
  if (le != hipSuccess) {
    (void)hipStreamSynchronize(st);
    throw HscError{HSC_ERR_HIP, std::string("PCG launch: ") + hipGetErrorString(le)};
  }
  if (after) after();
  std::vector<hsc::SysOut> out = collect_pcg(sc, nsys, st, fin ? hs : nullptr, fin);
  for (int s = 0; s < nsys; ++s) hs[s]->pairs_used = out[s].pairs != 0;
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  if (pcg_ms) *pcg_ms = ms;
  if (P.prof && max_iter_override <= 0) {
    std::vector<long long> pr((size_t)32 * nwg);
    HIPCHK(hipMemcpy(pr.data(), P.prof, sizeof(long long) * pr.size(), hipMemcpyDeviceToHost));
    const char* names[8] = {"init", "neumann", "A", "B", "T1", "T2", "", "nbr"};
    std::fprintf(stderr,
                 "[hsc stage timing] G=%d nwg=%d nL=%d iters=%d pcg_ms=%.3f (us per visit of each barrier site: "
                 "compute min/median/max over workgroups [argmax, LR?] | wait median)\n",
                 plan.G, nwg, P.n_lr_wg, out[0].iterations, ms);
    for (int si = 0; si < 8; ++si) {
      if (si == 6) continue;  // (slot 12 holds the stage-B reduction ticks)
      // workgroups that visited the site (the point-to-point pass sync
      // is visited by the row workgroups only)
      long long n = 0;
      std::vector<std::pair<double, int>> c;
      std::vector<double> w;
      for (int b = 0; b < nwg; ++b) {
        const long long nb = pr[(size_t)32 * b + 20 + si];
        if (!nb) continue;
        n = std::max(n, nb);
        c.push_back({pr[(size_t)32 * b + 2 * si] / 100.0 / nb, b});
        w.push_back(pr[(size_t)32 * b + 2 * si + 1] / 100.0 / nb);
      }
      if (!n) continue;
      std::sort(c.begin(), c.end());
      std::sort(w.begin(), w.end());
      const int am = c.back().second, mid = (int)c.size() / 2;
      std::fprintf(stderr, "  %-8s n=%6lld  %8.2f %8.2f %8.2f [wg%d, %s] | %8.2f\n", names[si], n, c.front().first,
                   c[mid].first, c.back().first, am, am < P.n_lr_wg ? "LR" : "row", w[mid]);
    }
    {
      // stage B split: reductions (slot 12) vs node tiles (rest of site 3)
      const long long n = pr[23];
      std::vector<double> red(nwg);
      for (int b = 0; b < nwg; ++b) red[b] = n ? pr[(size_t)32 * b + 12] / 100.0 / n : 0.0;
      std::sort(red.begin(), red.end());
      if (n) std::fprintf(stderr, "  B reductions part: min/median/max %8.2f %8.2f %8.2f\n", red.front(), red[nwg / 2], red.back());
    }
  }
  return out;
}

TuneKey tune_key(hsc_handle* const* hs, int nsys) {
  long long nbt = 0, nnz = 0;
  int kmax = 0;
  for (int s = 0; s < nsys; ++s) {
    nbt += hs[s]->nb;
    nnz += hs[s]->nnzb;
    kmax = std::max(kmax, hs[s]->k);
  }
  const hsc_config& c = hs[0]->cfg;
  const bool mixed = c.precision == HSC_PRECISION_MIXED;
  return TuneKey{device_info(hs[0]->device).cu,
                 (int)std::log2((double)std::max(nbt, 1LL)),
                 (int)(nnz / std::max(nbt, 1LL)),
                 kmax / 16,
                 mixed ? 1 : 0,
                 nsys > 1 ? 1 : 0,
                 c.preconditioner == HSC_PRECOND_NEUMANN ? 0 : 1,
                 c.neumann_order,
                 c.pass_sync | (c.pipelined << 2)};
}

bool pinned(const hsc_config& c) { return c.lanes_per_row && c.wg_per_cu && c.schedule; }

// The table's entry for the class, with the config's pinned fields applied;
// false when the class has no entry (and the config pins not everything).
bool tune_lookup(hsc_handle* const* hs, int nsys, LaunchCfg* out) {
  const hsc_config& c = hs[0]->cfg;
  if (pinned(c)) {
    *out = {c.lanes_per_row, c.wg_per_cu, c.schedule};
    return true;
  }
  std::lock_guard<std::mutex> g(g_tune_mu);
  load_tune_locked();
  auto it = g_tune.find(tune_key(hs, nsys));
  if (it == g_tune.end()) return false;
  *out = {c.lanes_per_row ? c.lanes_per_row : it->second.G, c.wg_per_cu ? c.wg_per_cu : it->second.wgpcu,
          c.schedule ? c.schedule : it->second.sched};
  return true;
}

// Time a few PCG iterations per candidate launch shape and keep the fastest
// in the table (and its file).  The systems' rhs, r and x must hold a
// right-hand side the timing runs may start from (they are overwritten).
LaunchCfg tune_now(hsc_handle* const* hs, int nsys, hipStream_t st) {
  const hsc_config& c = hs[0]->cfg;
  const int Gs[3] = {4, 8, 16};
  double best = 1e300;
  LaunchCfg bestc{8, 1, 1};
  for (int gi = 0; gi < 3; ++gi) {
    if (c.lanes_per_row && Gs[gi] != c.lanes_per_row) continue;
    for (int w = 1; w <= 2; ++w) {
      if (c.wg_per_cu && w != c.wg_per_cu) continue;
      for (int sc = 1; sc <= 3; ++sc) {
        if (c.schedule && sc != c.schedule) continue;
        // schedule 3 (the latency schedule on body-aligned row tiles) is a
        // candidate of its own at 4 lanes per row when the tiles differ from
        // the fixed ones
        if (sc == 3 && !c.schedule) {
          bool differs = Gs[gi] == 4;
          for (int s = 0; s < nsys && differs; ++s)
            differs = hs[s]->ntiles > 0 && hs[s]->ntiles != (hs[s]->nb + 127) / 128;
          if (!differs) continue;
        }
        LaunchCfg cand{Gs[gi], w, sc};
        try {
          float ms = 0.f;
          std::vector<hsc::SysOut> o;
          o = run_pcg(hs, nsys, cand, 12, st, &ms);  // warm
          o = run_pcg(hs, nsys, cand, 12, st, &ms);
          int it = 1;
          for (auto& x : o) it = std::max(it, x.iterations);
          const double per = ms / it;
          if (per < best) {
            best = per;
            bestc = cand;
          }
        } catch (const HscError& e) {
          if (e.code == HSC_ERR_CAPACITY) continue;  // candidate does not fit: try the next
          throw;                                     // a HIP or device error ends the tuning
        }
      }
    }
  }
  if (!pinned(c)) {
    const TuneKey key = tune_key(hs, nsys);
    const TuneVal v{bestc.G, bestc.wgpcu, bestc.sched};
    std::lock_guard<std::mutex> g(g_tune_mu);
    load_tune_locked();
    g_tune[key] = v;
    save_tune_locked(key, v);
  }
  return bestc;
}

// Launch shape and scratch of a handle's persistent PCG, fixed at setup so
// that a solve allocates nothing and runs no tuning: the shape comes from the
// tune table, or is timed now (from a pseudo-random right-hand side, with
// the handle's spare vector as x) when the class is new; then every buffer of
// that launch is sized by planning it once.
void prepare_launch(hsc_handle* h, hipStream_t st, bool retune = false) {
  hsc_handle* hs[1] = {h};
  h->launch_ready = false;
  h->xsol = h->yfull.as<double>();
  LaunchCfg lc{};
  if (retune || !tune_lookup(hs, 1, &lc)) {
    HIPCHK(hsc::launch_hash_fill(h->rhs.as<double>(), h->fn, 0x2545f491u, st));
    HIPCHK(hipMemcpyAsync(h->r.p, h->rhs.p, sizeof(double) * h->fn, hipMemcpyDeviceToDevice, st));
    lc = tune_now(hs, 1, st);
  }
  // plan with the last solve's x: a solve into the same buffer then finds
  // its descriptor already uploaded (nothing runs with it before a solve,
  // which plans with its own x)
  if (h->last_xsol) h->xsol = h->last_xsol;
  {
    PcgPlan plan = plan_pcg(hs, 1, lc, 0, st, h->pcg);
    ensure_host_scratch(h->pcg, 1);
    static const bool stage_timing = std::getenv("HSC_STAGE_TIMING") != nullptr;
    if (stage_timing) HIPCHK(h->profbuf.ensure(sizeof(long long) * 32 * plan.P.nwg));
  }
  h->xsol = nullptr;
  h->G = lc.G;
  h->wgpcu = lc.wgpcu;
  h->sched = lc.sched;
  h->launch_ready = true;
}

// ------------------------------------------------------------------ setup
// Per-pass preconditioner weights.  Neumann: omega for every pass.
// Chebyshev: the passes z0 = w0 Dinv r, z_j = z_(j-1) + w_j Dinv (r - D z_(j-1))
// have residual polynomial prod_j (1 - w_j mu) over the eigenvalues mu of
// Dinv D; w_j = 1 / mu_j with mu_j the roots of the degree-m Chebyshev
// polynomial shifted to [a, b] makes it the Chebyshev polynomial of that
// interval (the min-max choice).  b = 1.2 x a 12-step power-iteration
// estimate of lambda_max (from below), a = b / poly_interval_ratio.  The
// preconditioner stays SPD for every mu in (0, b]: there each factor
// 1 - mu / mu_j of the residual polynomial keeps the product below 1.
constexpr int kPowerSteps = 12;

// Device buffer of the power iteration: x, z (n each), two partial arrays,
// the result |Dinv D x_11|^2.
size_t power_buffer_doubles(int n, int nb) {
  const size_t cap = (size_t)std::max(hsc::power_step_tiles(8, nb), 256);
  return 2 * (size_t)n + 2 * cap + 1;
}

// Enqueue the power iteration for lambda_max(Dinv D) on st (D's rows and
// inverse diagonal blocks ready there): 12 fused steps from a fixed
// pseudo-random start, each normalising its input by the fixed-order norm of
// the previous step's partials; returns the device address that holds
// |Dinv D x_11|^2 with |x_11| = 1 once st has drained.
const double* enqueue_power_iteration(hsc_handle* h, hipStream_t st) {
  const int n = h->fn, nb = h->nb;
  const size_t cap = (size_t)std::max(hsc::power_step_tiles(8, nb), 256);
  double* x = h->pwr.as<double>();
  double* z = x + n;
  double* pa = z + n;
  double* pb = pa + cap;
  double* lam2 = pb + cap;
  hsc::SysDesc S = make_desc(h);
  HIPCHK(hsc::launch_hash_fill(x, n, 0x9e3779b9u, st));
  int np_in = 0;
  HIPCHK(hsc::launch_dot_partial(x, x, n, pa, &np_in, st));  // |x_0|^2 partials
  double* bufs[2] = {x, z};
  double* parts[2] = {pa, pb};
  const int tiles = hsc::power_step_tiles(8, nb);
  for (int it = 0; it < kPowerSteps; ++it) {
    HIPCHK(hsc::launch_power_step(8, S, bufs[it & 1], parts[it & 1], np_in, bufs[(it + 1) & 1],
                                  parts[(it + 1) & 1], st));
    np_in = tiles;
  }
  HIPCHK(hsc::launch_sum_partials(parts[kPowerSteps & 1], tiles, lam2, st));
  return lam2;
}

int interval_from_prediction(hsc_handle* h, const double pred[3], double* ratio);
const double* enqueue_prediction(hsc_handle* h, hipStream_t st);

// Per-pass weights: omega for the Neumann series (and for the learned
// preconditioner when its network's output is rejected); for the Chebyshev
// and learned preconditioners the reciprocal Chebyshev roots of [b / ratio,
// b] with b = 1.2 x sqrt(lam2) (see the comment above), the ratio from the
// config (Chebyshev) or from the network's node means `pred` (learned).
void set_poly_weights(hsc_handle* h, double lam2, const double* pred) {
  const int m = h->cfg.neumann_order;
  for (int j = 0; j < hsc::kMaxPolySteps; ++j) h->omegas[j] = h->cfg.omega;
  h->lambda_max = 0.0;
  h->poly_ratio = 0.0;
  h->learned_used = 0;
  h->fallback_reason = 0;
  if (h->cfg.preconditioner == HSC_PRECOND_NEUMANN) return;
  const double lam = std::sqrt(lam2);
  if (!(lam > 0) || !std::isfinite(lam)) throw HscError{HSC_ERR_NONFINITE_FACTORS, "HSC: spectrum estimate failed"};
  h->lambda_max = lam;
  double ratio = h->cfg.poly_interval_ratio;
  if (h->cfg.preconditioner == HSC_PRECOND_LEARNED) {
    const int why = interval_from_prediction(h, pred, &ratio);
    if (why) {
      // the plain Neumann series, its weight capped at 2 / b so that every
      // factor 1 - w mu stays in [-1, 1) on (0, b]: SPD at any order
      h->fallback_reason = why;
      const double w = std::min(h->cfg.omega, 2.0 / (1.2 * lam));
      for (int j = 0; j < hsc::kMaxPolySteps; ++j) h->omegas[j] = w;
      return;
    }
    h->learned_used = 1;
  }
  h->poly_ratio = ratio;
  const double b = 1.2 * lam, a = b / ratio;
  std::vector<double> mu(m);
  for (int j = 0; j < m; ++j) mu[j] = 0.5 * (b + a) + 0.5 * (b - a) * std::cos((2.0 * j + 1.0) * M_PI / (2.0 * m));
  std::sort(mu.begin(), mu.end());
  // Leja order: the largest root first, then repeatedly the root maximising
  // the product of distances to those taken.  Every partial product of the
  // passes' factors (1 - mu / mu_j), and every remaining one, stays bounded
  // (about 1e2 at degree 16), where ascending roots grow it to 1e7 and FP32
  // passes lose all precision beyond degree ~16.
  std::vector<double> ord;
  ord.push_back(mu.back());
  mu.pop_back();
  while (!mu.empty()) {
    size_t best = 0;
    double best_v = -INFINITY;
    for (size_t i = 0; i < mu.size(); ++i) {
      double v = 0.0;
      for (double o : ord) v += std::log(std::fabs(mu[i] - o));
      if (v > best_v) {
        best = i;
        best_v = v;
      }
    }
    ord.push_back(mu[best]);
    mu.erase(mu.begin() + (long)best);
  }
  for (int j = 0; j < m; ++j) h->omegas[j] = 1.0 / ord[j];
  for (int j = m; j < hsc::kMaxPolySteps; ++j) h->omegas[j] = h->omegas[m - 1];
}

// Small setup results fetched from the device (pinned host memory).
struct SetupStage {
  int herr[8];
  int nnzb, nact;
  int ints2[2];
  int aca_err, max_coupled;
  int tiles[2];  // body-aligned row tiles: count, most coupled nodes in one
  double dm[4];
  double lam2, fst;
  double pred[4];  // learned: the network's node-mean outputs
  double hist[hsc::kMaxRankLds + 1];
};

// Synchronises the setup's second stream when the setup leaves (normally or
// by an exception), so no D-side kernel outlives the call.
struct AuxJoin {
  hipStream_t s;
  ~AuxJoin() {
    if (s) (void)hipStreamSynchronize(s);
  }
};

void do_setup(hsc_handle* h, const hsc_bcoo3_view* Av, int qo, int qm, int fo, int fn, hipStream_t st) {
  auto t0 = Clock::now();
  h->setup_ok = false;
  h->launch_ready = false;
  ++h->gen;
  for (bool& ok : h->tnbr_ok) ok = false;
  h->pairs_hbl = -1;
  h->strip_nch = -1;
  h->ntiles = 0;
  if (!Av) throw HscError{HSC_ERR_INVALID_ARGUMENT, "null matrix view"};
  const long long N = 3 * Av->block_dim;
  // linear_hsc.cu:197-200
  if (!(qm > 0 && fn > 0 && qm % 12 == 0 && fn % 3 == 0 && (long long)qm + fn == N && qo % 3 == 0 && fo % 3 == 0 &&
        qo >= 0 && fo >= 0 && qo + qm <= N && fo + fn <= N && (qo + qm <= fo || fo + fn <= qo)))
    throw HscError{HSC_ERR_INVALID_ARGUMENT, "HSC requires only ABD (12/body) and FEM (3/node) DOFs"};
  if (Av->triplet_count < 0 || (Av->triplet_count > 0 && (!Av->row_indices || !Av->col_indices || !Av->values)))
    throw HscError{HSC_ERR_INVALID_ARGUMENT, "invalid BCOO view"};
  if (Av->triplet_count >= (1LL << 30)) throw HscError{HSC_ERR_CAPACITY, "too many triplets"};
  h->N = (int)N;
  h->qo = qo;
  h->qm = qm;
  h->fo = fo;
  h->fn = fn;
  h->nb = fn / 3;
  h->count = Av->triplet_count;
  const int nb = h->nb;
  HIPCHK(h->row_count.ensure(sizeof(int) * nb));
  HIPCHK(h->rowptr.ensure(sizeof(int) * (nb + 1)));
  HIPCHK(h->cursor.ensure(sizeof(int) * nb));
  HIPCHK(h->active_flag.ensure(sizeof(int) * nb));
  HIPCHK(h->act_prefix.ensure(sizeof(int) * (nb + 1)));
  HIPCHK(h->act_node.ensure(sizeof(int) * nb));
  HIPCHK(h->invD64.ensure(sizeof(double) * 9 * nb));
  HIPCHK(h->invD32.ensure(sizeof(float) * 9 * nb));
  HIPCHK(h->A.ensure(sizeof(double) * qm * qm));
  HIPCHK(h->ints.ensure(sizeof(int) * 32));
  HIPCHK(h->stage.ensure(sizeof(SetupStage)));
  HIPCHK(h->stage_dev.ensure(sizeof(SetupStage)));
  SetupStage* ss = h->stage.as<SetupStage>();
  SetupStage* sd = h->stage_dev.as<SetupStage>();
  // gather device words into the device copy of the stage (then one copy
  // of a field range brings them to the pinned one)
  hsc::GatherList gl{};
  auto gather = [&](void* dst_field, const void* src, size_t bytes) {
    if (gl.count >= hsc::kGatherMax) throw HscError{HSC_ERR_HIP, "setup: gather list full"};
    gl.dst[gl.count] = static_cast<unsigned*>(dst_field);
    gl.src[gl.count] = static_cast<const unsigned*>(src);
    gl.n[gl.count] = (int)(bytes / 4);
    ++gl.count;
  };
  auto fetch = [&](const void* first_field, size_t bytes) {
    const size_t off = (size_t)(static_cast<const char*>(first_field) - reinterpret_cast<const char*>(sd));
    HIPCHK(hsc::launch_gather_words(gl, st));
    HIPCHK(hipMemcpyAsync(reinterpret_cast<char*>(ss) + off, first_field, bytes, hipMemcpyDeviceToHost, st));
    gl.count = 0;
  };
  {
    const void* zp[] = {h->row_count.p, h->cursor.p, h->active_flag.p, h->invD64.p, h->invD32.p, h->A.p, h->ints.p};
    const size_t zb[] = {sizeof(int) * nb, sizeof(int) * nb, sizeof(int) * nb, sizeof(double) * 9 * nb,
                         sizeof(float) * 9 * nb, sizeof(double) * qm * qm, sizeof(int) * 32};
    HIPCHK(hsc::launch_zero_multi(zp, zb, 7, st));
  }
  int* err = h->ints.as<int>();
  hsc::SplitParams sp{Av->row_indices, Av->col_indices, Av->values, Av->triplet_count, qo, qm, fo, fn,
                      h->row_count.as<int>(), h->active_flag.as<int>(), h->invD64.as<double>(),
                      h->invD32.as<float>(), h->A.as<double>(), err};
  HIPCHK(hsc::launch_split(sp, st));
  HIPCHK(hsc::launch_scan2(h->row_count.as<int>(), h->rowptr.as<int>(), h->active_flag.as<int>(),
                           h->act_prefix.as<int>(), nb, st));
  gather(sd->herr, err, sizeof(sd->herr));
  gather(&sd->nnzb, h->rowptr.as<int>() + nb, sizeof(int));
  gather(&sd->nact, h->act_prefix.as<int>() + nb, sizeof(int));
  static_assert(offsetof(SetupStage, nact) == offsetof(SetupStage, herr) + sizeof(int) * 9, "stage layout");
  fetch(sd->herr, offsetof(SetupStage, nact) + sizeof(int) - offsetof(SetupStage, herr));
  HIPCHK(hipStreamSynchronize(st));
  const int* herr = ss->herr;
  const int nnzb = ss->nnzb, nact = ss->nact;
  if (herr[1]) throw HscError{HSC_ERR_INVALID_ARGUMENT, "HSC requires only ABD (12/body) and FEM (3/node) DOFs"};
  if (herr[3]) throw HscError{HSC_ERR_UNEXPECTED_COUPLING, "Unexpected HSC subsystem coupling"};
  if (herr[4]) throw HscError{HSC_ERR_NONFINITE_FACTORS, "HSC: non-finite Hessian entries"};
  if (herr[0])
    throw HscError{HSC_ERR_FEM_DIAGONAL, "HSC: " + std::to_string(herr[0]) + " non-positive FEM diagonal blocks"};
  h->blockdiag = herr[2] == 0;
  h->nnzb = nnzb;
  h->a = 3 * nact;
  const int a = h->a;
  const int limit = std::min(std::min(h->cfg.max_rank, qm), a);
  h->limit = limit;
  const int lim1 = std::max(limit, 1);
  HIPCHK(h->act_dof.ensure(sizeof(int) * std::max(a, 1)));
  HIPCHK(h->key_col.ensure(sizeof(int) * nnzb));
  HIPCHK(h->key_src.ensure(sizeof(int) * nnzb));
  HIPCHK(h->colidx.ensure(sizeof(int) * nnzb));
  HIPCHK(h->val64.ensure(sizeof(double) * hsc::val_count(nnzb)));
  HIPCHK(h->val32.ensure(sizeof(float) * hsc::val_count(nnzb)));
  const int nch = (nnzb + 63) / 64;
  HIPCHK(h->hbits.ensure(sizeof(unsigned long long) * (size_t)std::max(nch, 1)));
  HIPCHK(h->hrow.ensure(sizeof(int) * (size_t)(nch + 1)));
  HIPCHK(h->tstart.ensure(sizeof(int) * (size_t)hsc::row_tiles_capacity()));
  HIPCHK(h->reach.ensure(sizeof(int) * (size_t)nb));
  HIPCHK(h->E.ensure(sizeof(double) * (size_t)qm * std::max(a, 1)));
  HIPCHK(h->R0.ensure(sizeof(double) * (size_t)qm * std::max(a, 1)));
  HIPCHK(h->R1.ensure(sizeof(double) * (size_t)qm * std::max(a, 1)));
  HIPCHK(h->U.ensure(sizeof(double) * (size_t)qm * lim1));
  HIPCHK(h->V.ensure(sizeof(double) * (size_t)std::max(a, 1) * lim1));
  HIPCHK(h->hist.ensure(sizeof(double) * (lim1 + 1)));
  HIPCHK(h->pivots.ensure(sizeof(int) * 2 * lim1));
  HIPCHK(h->Binv.ensure(sizeof(double) * 144 * (qm / 12)));
  HIPCHK(h->P.ensure(sizeof(double) * qm * qm));
  HIPCHK(h->s1.ensure(sizeof(double) * qm * qm));
  HIPCHK(h->s2.ensure(sizeof(double) * qm * qm));
  HIPCHK(h->s3.ensure(sizeof(double) * qm * qm));
  HIPCHK(h->W1.ensure(sizeof(double) * 2 * (size_t)qm * lim1));  // P U, then (P U)^T
  HIPCHK(h->C.ensure(sizeof(double) * (size_t)lim1 * lim1));
  HIPCHK(h->Lm.ensure(sizeof(double) * (size_t)lim1 * lim1));
  HIPCHK(h->W64.ensure(sizeof(double) * (size_t)std::max(a, 1) * lim1));
  HIPCHK(h->W32.ensure(sizeof(float) * (size_t)std::max(a, 1) * lim1));
  for (Buf* b : {&h->rhs, &h->r, &h->p0, &h->p1, &h->z0, &h->z1, &h->q, &h->q64, &h->yfull})
    HIPCHK(b->ensure(sizeof(double) * fn));
  HIPCHK(h->small.ensure(sizeof(double) * (qm + 2 * lim1 + 64)));
  HIPCHK(h->ybuf.ensure(sizeof(double) * lim1));
  HIPCHK(h->misc.ensure(sizeof(double) * 64));
  HIPCHK(h->respart.ensure(sizeof(double) * 3 * ((fn + hsc::kBlock - 1) / hsc::kBlock + qm + 1)));
  // the polynomial preconditioners (Chebyshev, learned) need the spectrum bound
  const bool cheb = h->cfg.preconditioner != HSC_PRECOND_NEUMANN;
  if (cheb) HIPCHK(h->pwr.ensure(sizeof(double) * power_buffer_doubles(fn, nb)));
  if (!h->aux) {
    HIPCHK(hipStreamCreateWithFlags(&h->aux, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&h->aux_ev, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&h->aux_done, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&h->sort_ev, hipEventDisableTiming));
  }
  const bool learned = h->cfg.preconditioner == HSC_PRECOND_LEARNED;
  if (learned && !h->aux2) {
    HIPCHK(hipStreamCreateWithFlags(&h->aux2, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&h->d_ev, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&h->aux2_done, hipEventDisableTiming));
  }
  HIPCHK(h->barrier.ensure(sizeof(unsigned) * hsc::kBarrierWords));
  HIPCHK(h->errflag.ensure(64));
  {
    // E (filled by k_place), the ACA history, grid-barrier counters and error flag
    const void* zp[] = {h->hist.p, h->barrier.p, h->errflag.p, h->E.p};
    const size_t zb[] = {sizeof(double) * (lim1 + 1), sizeof(unsigned) * hsc::kBarrierWords, 64,
                         sizeof(double) * (size_t)qm * a};
    HIPCHK(hsc::launch_zero_multi(zp, zb, a > 0 ? 4 : 3, st));
  }
  HIPCHK(hsc::launch_active(h->active_flag.as<int>(), h->act_prefix.as<int>(), nb, h->act_node.as<int>(),
                            h->act_dof.as<int>(), st));
  // coupled nodes per 128-node tile (err[12]): sizes the LDS cache of W rows
  // used by the persistent PCG's resident node tiles
  HIPCHK(hsc::launch_max_tile_active(h->act_node.as<int>(), nb, 128, err + 12, st));
  hsc::PlaceParams pp{Av->row_indices, Av->col_indices, Av->values, Av->triplet_count, qo, qm, fo, fn,
                      h->rowptr.as<int>(), h->cursor.as<int>(), h->key_col.as<int>(), h->key_src.as<int>(),
                      h->act_node.as<int>(), a, h->E.as<double>()};
  HIPCHK(hsc::launch_place(pp, st));
  // the D side waits for this point only
  HIPCHK(hipEventRecord(h->aux_ev, st));
  AuxJoin join{nullptr};
  // ---- ACA
  int* rank_dev = err + 8;
  double* dmisc = h->misc.as<double>();
  if (limit > 0) {
    const DeviceInfo di = device_info(h->device);
    const long long total = (long long)qm * a;
    // register-resident residual when it fits in one workgroup per CU,
    // else the residual streamed through memory each step
    int occ_reg = 0;
    HIPCHK(hsc::aca_reg_occupancy(&occ_reg));
    const int nwg_reg = occ_reg > 0 ? hsc::aca_reg_grid(qm, a, di.cu) : 0;
    int nwg = nwg_reg;
    if (nwg == 0) {
      int occ = 0;
      HIPCHK(hsc::aca_occupancy(&occ));
      // one workgroup per CU at most: each ACA step is one grid barrier plus
      // a pass over E, so fewer, fuller workgroups are cheaper
      nwg = (int)std::min<long long>(di.cu, std::max<long long>(1, (total + 2047) / 2048));
      nwg = std::max(1, std::min(nwg, std::max(occ, 1) * di.cu));
    }
    HIPCHK(h->acapart.ensure(sizeof(double) * 8 * nwg));  // two halves, see k_aca
    if (nwg_reg) HIPCHK(h->acacand.ensure(sizeof(double) * 2 * (size_t)nwg * qm));
    hsc::AcaParams ap{};
    ap.E = h->E.as<double>();
    ap.R0 = h->R0.as<double>();
    ap.R1 = h->R1.as<double>();
    ap.qm = qm;
    ap.a = a;
    ap.limit = limit;
    ap.stop = h->cfg.aca_rtol;
    ap.U = h->U.as<double>();
    ap.V = h->V.as<double>();
    ap.hist = h->hist.as<double>();
    ap.pivots = h->pivots.as<int>();
    ap.part = h->acapart.as<double>();
    ap.cand = nwg_reg ? h->acacand.as<double>() : nullptr;
    ap.rank_out = rank_dev;
    ap.scale_out = dmisc + 10;
    ap.barrier = h->barrier.as<unsigned>();
    ap.error = h->errflag.as<int>();
    ap.timeout_ticks = kTimeoutTicks;
    ap.nwg = nwg;
    // ACA by ABD body first when E is block diagonal by body (each block by
    // one workgroup, no grid barrier); the whole-E kernel then exits at once
    const int nbody = qm / 12, nbb = std::min(nbody, hsc::kAcaBlkMaxBodies);
    HIPCHK(h->acab.ensure(sizeof(double) * ((size_t)a + 1 + 157 * (size_t)nbb)));
    HIPCHK(h->acabi.ensure(sizeof(int) * ((size_t)a + 4 + nbody + 25 * (size_t)nbb)));
    hsc::AcaBlkParams bp{};
    bp.E = h->E.as<double>();
    bp.qm = qm;
    bp.a = a;
    bp.limit = limit;
    bp.stop = h->cfg.aca_rtol;
    bp.colsq = h->acab.as<double>();
    bp.scale = bp.colsq + a;
    bp.ublk = bp.scale + 1;
    bp.hblk = bp.ublk + 144 * (size_t)nbb;
    bp.vtmp = h->R0.as<double>();  // a x 12 <= qm x a
    bp.colbody = h->acabi.as<int>();
    bp.plan = bp.colbody + a;
    bp.pblk = bp.plan + 4 + nbody;
    bp.rank_b = bp.pblk + 24 * (size_t)nbb;
    bp.U = ap.U;
    bp.V = ap.V;
    bp.hist = ap.hist;
    bp.pivots = ap.pivots;
    bp.rank_out = ap.rank_out;
    bp.scale_out = ap.scale_out;
    HIPCHK(hsc::launch_aca_blocks(bp, st));
    ap.skip = bp.plan;
    HIPCHK(nwg_reg ? hsc::launch_aca_reg(ap, st) : hsc::launch_aca(ap, st));
  }
  // ---- ABD inverse and P
  HIPCHK(hsc::launch_body_inverse(h->A.as<double>(), qm, h->Binv.as<double>(), err + 9, st));
  hsc::AbdParams abp{h->A.as<double>(), h->Binv.as<double>(), h->P.as<double>(), h->s1.as<double>(),
                     h->s2.as<double>(), h->s3.as<double>(), qm, h->cfg.abd_inverse, h->blockdiag ? 1 : 0,
                     h->cfg.neumann_order, h->cfg.omega, h->cfg.abd_inverse_tol, dmisc};
  if (h->cfg.abd_inverse == HSC_ABD_NEUMANN && h->cfg.adaptive_abd_inverse) abp.mode = 2;
  HIPCHK(hsc::launch_abd_inverse(abp, st));
  h->abd_mode = abp.mode;
  // ---- C = U^T P U, its Cholesky C = L L^T and W = V L, for the rank ACA
  // found: the kernels read it on the device, so the chain follows ACA
  // without a round trip to the host (grids sized for the largest rank)
  if (limit > 0) {
    hsc::FactorParams fp{h->P.as<double>(), h->U.as<double>(), h->V.as<double>(), qm, a, limit, rank_dev,
                         h->W1.as<double>(), h->W1.as<double>() + (size_t)qm * limit, h->C.as<double>(),
                         h->Lm.as<double>(), dmisc + 4};
    HIPCHK(hsc::launch_coupling_core(fp, st));
    HIPCHK(hsc::launch_make_W(h->V.as<double>(), h->Lm.as<double>(), a, rank_dev, limit, h->W64.as<double>(),
                              h->W32.as<float>(), st));
  }
  // The D side (row sort and value gather, then the network and the power
  // iteration) depends on nothing after k_place.  It is submitted once the
  // main stream's short ACA / ABD-inverse chain and its read-backs are
  // queued, so that the host's launch time for the two longer chains
  // overlaps that chain's execution and the ACA kernels start first.
  HIPCHK(hipStreamWaitEvent(h->aux, h->aux_ev, 0));
  join.s = h->aux;
  HIPCHK(hsc::launch_sort_gather(h->rowptr.as<int>(), nb, h->key_col.as<int>(), h->key_src.as<int>(), Av->values,
                                 nnzb, h->colidx.as<int>(), h->val64.as<double>(), h->val32.as<float>(), h->aux));
  HIPCHK(hsc::launch_chunk_heads(h->rowptr.as<int>(), nb, nnzb, h->hbits.as<unsigned long long>(),
                                 h->hrow.as<int>(), h->aux));
  h->strip_nch = nch;
  HIPCHK(hipEventRecord(h->sort_ev, h->aux));
  AuxJoin join2{nullptr};
  // the network needs D only: it runs on a third stream beside the power
  // iteration, enqueued before it (its GPU chain is the longer one: started
  // first, the two end together)
  if (learned) {
    HIPCHK(hipEventRecord(h->d_ev, h->aux));
    HIPCHK(hipStreamWaitEvent(h->aux2, h->d_ev, 0));
    join2.s = h->aux2;
  }
  const double* pred_dev = learned ? enqueue_prediction(h, h->aux2) : nullptr;
  const double* lam2_dev = cheb ? enqueue_power_iteration(h, h->aux) : nullptr;
  // body-aligned row tiles from D's columns, on the main stream once the
  // column sort is done (it has slack after the coupling core there),
  // enqueued after the two longer D-side chains
  HIPCHK(hipStreamWaitEvent(st, h->sort_ev, 0));
  HIPCHK(hsc::launch_row_tiles(h->rowptr.as<int>(), h->colidx.as<int>(), h->act_node.as<int>(), nb,
                               h->reach.as<int>(), h->tstart.as<int>(), err + 16, st));
  // the main stream waits for both D-side chains, so that one
  // synchronisation returns every result of the three streams
  HIPCHK(hipEventRecord(h->aux_done, h->aux));
  HIPCHK(hipStreamWaitEvent(st, h->aux_done, 0));
  if (learned) {
    HIPCHK(hipEventRecord(h->aux2_done, h->aux2));
    HIPCHK(hipStreamWaitEvent(st, h->aux2_done, 0));
  }
  // every small result of the three streams by one gather and one copy
  // (zeros where a stage did not run)
  gather(sd->ints2, err + 8, sizeof(sd->ints2));
  gather(&sd->aca_err, limit > 0 ? h->errflag.p : nullptr, sizeof(int));
  gather(&sd->max_coupled, err + 12, sizeof(int));
  gather(sd->tiles, err + 16, sizeof(sd->tiles));
  gather(sd->dm, dmisc, sizeof(sd->dm));
  gather(&sd->lam2, lam2_dev, sizeof(double));
  gather(&sd->fst, limit > 0 ? dmisc + 4 : nullptr, sizeof(double));
  gather(sd->pred, pred_dev, sizeof(double) * 3);
  gather(sd->pred + 3, nullptr, sizeof(double));
  gather(sd->hist, h->hist.as<double>(), sizeof(double) * (lim1 + 1));
  static_assert(offsetof(SetupStage, hist) - offsetof(SetupStage, ints2) ==
                    sizeof(int) * 6 + sizeof(double) * 10,
                "stage layout: the gathered fields are contiguous");
  fetch(sd->ints2, offsetof(SetupStage, hist) - offsetof(SetupStage, ints2) + sizeof(double) * (lim1 + 1));
  HIPCHK(hipStreamSynchronize(st));
  join.s = join2.s = nullptr;  // both chains have ended
  const int* ints2 = ss->ints2;
  const double* dm = ss->dm;
  const int aca_err = ss->aca_err;
  h->max_coupled128 = ss->max_coupled;
  h->ntiles = ss->tiles[0];
  h->max_coupled_tiles = ss->tiles[1];
  if (aca_err) throw HscError{HSC_ERR_DEVICE_TIMEOUT, "ACA: bounded grid wait expired"};
  const int rank = limit > 0 ? ints2[0] : 0;
  if (ints2[1]) throw HscError{HSC_ERR_ABD_NOT_SPD, "HSC: non-SPD ABD block"};
  const int abd_status = (int)dm[2];
  h->abd_res = dm[0];
  h->doublings = (int)dm[1];
  if (abd_status == 4) throw HscError{HSC_ERR_ABD_NOT_SPD, "HSC: non-SPD ABD block"};
  if (abd_status == 5)
    throw HscError{HSC_ERR_ABD_INVERSE, "HSC ABD inverse residual exceeds tolerance after " +
                                            std::to_string(h->doublings) + " doublings"};
  if (abd_status == 6) throw HscError{HSC_ERR_NONFINITE_FACTORS, "HSC: nonfinite Neumann/Schur factors"};
  h->aca_err = (limit > 0) ? ss->hist[rank] : 0.0;
  h->k = rank;
  if (rank > 0 && ss->fst != 0.0) throw HscError{HSC_ERR_NONFINITE_FACTORS, "HSC: nonfinite or non-SPD coupling core C"};
  set_poly_weights(h, ss->lam2, ss->pred);
  prepare_launch(h, st);
  h->setup_ok = true;
  h->setup_ms = ms_since(t0);
}

// ------------------------------------------------------------------ learned polynomial
// The learned preconditioner keeps the Chebyshev-weighted block-Jacobi
// passes (a fixed SPD polynomial of Dinv D, so plain PCG applies) and lets a
// graph MLP choose the polynomial's interval: at setup the network reads
// per-node features of D, its 3 outputs are averaged over the nodes, and the
// mean predicts log10 of the interval ratio b / a for degrees 8, 12 and 16
// (channel value = log10(ratio) - 2.5; other degrees interpolate between
// channels, beyond 16 the last one).  An output that is not finite or lies
// outside log10(ratio) in [1, 4.5] is rejected and the plain Neumann series
// (weight omega every pass) is used instead.
constexpr double kLearnedCenter = 2.5, kLearnedMinLog = 1.0, kLearnedMaxLog = 4.5;

void ensure_graph(hsc_handle* h, hipStream_t st) {
  const int nb = h->nb, ne = h->nnzb - h->nb;
  HIPCHK(h->nbr_ptr.ensure(sizeof(int) * (nb + 1)));
  HIPCHK(h->nbr.ensure(sizeof(int) * std::max(ne, 1)));
  HIPCHK(h->dfeat.ensure(sizeof(float) * 9 * nb));
  HIPCHK(h->sdiag.ensure(sizeof(float) * nb));
  HIPCHK(hsc::launch_node_graph(h->rowptr.as<int>(), h->colidx.as<int>(), h->val64.as<double>(), nb,
                                 h->nbr_ptr.as<int>(), h->nbr.as<int>(), h->dfeat.as<float>(), h->sdiag.as<float>(),
                                 st));
}

// The handle's network over its current node graph (the handle's own
// neighbour CSR, read in place; the net is reallocated only for a larger
// graph), with the loaded weights (cast to bf16 again only when they
// changed).
void ensure_mlp(hsc_handle* h, hipStream_t st) {
  if (h->learned_bytes != (long long)sizeof(float) * hsc::MlpLayout::count)
    throw HscError{HSC_ERR_INVALID_ARGUMENT, "learned preconditioner selected but no weights set"};
  ensure_graph(h, st);
  const int ne = h->nnzb - h->nb;
  if (!h->mlp || !h->mlp->fits(h->nb, ne)) {
    h->mlp.reset(new hsc::MlpNet());
    HIPCHK(h->mlp->init(h->nb, h->nbr_ptr.as<int>(), h->nbr.as<int>(), ne, 1, st));
    h->mlp_weights_version = -1;
  }
  h->mlp->borrow_graph(h->nb, h->nbr_ptr.as<int>(), h->nbr.as<int>());
  if (h->mlp_weights_version != h->learned_version) {
    HIPCHK(h->mlp->set_params(h->learned.as<float>(), true, st));
    h->mlp_weights_version = h->learned_version;
  }
  HIPCHK(h->xfeat.ensure(sizeof(hsc::bf16_t) * 16 * h->nb));
  HIPCHK(h->omlp.ensure(sizeof(float) * 3 * h->nb));
  HIPCHK(h->lpred.ensure(sizeof(double) * (4 + 3 * (size_t)hsc::mean_rows3_parts(h->nb))));
}

// The network's input features of h's system (graph and D's diagonal
// features built on st first).
void enqueue_learned_features(hsc_handle* h, hipStream_t st) {
  HIPCHK(hsc::launch_poly_features(h->rowptr.as<int>(), h->colidx.as<int>(), h->val64.as<double>(),
                                   h->dfeat.as<float>(), h->sdiag.as<float>(), h->act_node.as<int>(), h->nb,
                                   h->xfeat.as<hsc::bf16_t>(), st));
}

// Enqueue the network on h's system on st (behind D's rows there):
// features, forward, node mean of the 3 outputs; returns the device address
// of the 3 means (set once st drains).
const double* enqueue_prediction(hsc_handle* h, hipStream_t st) {
  ensure_mlp(h, st);
  enqueue_learned_features(h, st);
  HIPCHK(h->mlp->forward(h->xfeat.as<hsc::bf16_t>(), 1, h->omlp.as<float>(), st));
  double* lp = h->lpred.as<double>();
  HIPCHK(hsc::launch_mean_rows3(h->omlp.as<float>(), h->nb, lp + 4, lp, st));
  return lp;
}

// The interval ratio for h's degree from the network's 3 node means; returns
// 0 (accepted, *ratio set) or the fallback reason 2 (rejected).
int interval_from_prediction(hsc_handle* h, const double pred[3], double* ratio) {
  for (int c = 0; c < 3; ++c) h->learned_out[c] = pred[c];
  const double t = std::min(std::max((h->cfg.neumann_order - 8) / 4.0, 0.0), 2.0);
  const int c0 = (int)std::floor(t), c1 = std::min(c0 + 1, 2);
  const double u = (1.0 - (t - c0)) * pred[c0] + (t - c0) * pred[c1];
  const double lg = u + kLearnedCenter;
  if (!std::isfinite(lg) || lg < kLearnedMinLog || lg > kLearnedMaxLog) return 2;
  *ratio = std::pow(10.0, lg);
  return 0;
}


// Algorithmic HBM bytes of one Schur-PCG iteration of h's system (every
// matrix and vector element read or written once per pass).  FP64: an SpMV
// pass is 76 B per block (values + column), 4 B per row pointer and 16 B per
// DOF (one vector gathered, one written); each further preconditioner pass
// adds the inverse diagonal blocks (72 B per node) and r; stage A reads p_old
// and writes p; the low-rank apply reads W twice plus its indices; stage B
// reads q, p, r, x, the inverse blocks and writes r, x, z.  Mixed: 40 B per
// block, 36 B per inverse block, FP32 p, z, q (r and x stay FP64).
double pcg_bytes_per_iteration(const hsc_handle* h) {
  const double nb = h->nb, n = 3.0 * h->nb, nnzb = h->nnzb, a = h->k > 0 ? h->a : 0, k = h->k;
  const int order = h->cfg.neumann_order;
  if (h->cfg.precision != HSC_PRECISION_MIXED) {
    const double spmv = 76 * nnzb + 4 * (nb + 1) + 16 * n;
    return (order - 1) * (spmv + 72 * nb + 8 * n) + spmv + 16 * n + 2 * 8 * a * k + 4 * a + 6 * 8 * n + 72 * nb;
  }
  const double spmv = 40 * nnzb + 4 * (nb + 1) + 8 * n;
  return (order - 1) * (spmv + 36 * nb + 8 * n) + spmv + 8 * n + 2 * 4 * a * k + 4 * a + 4 * 8 * n + 2 * 4 * n +
         36 * nb + 4 * n;
}

constexpr double kHbmPeakGBs = 8000.0;  // MI355X HBM3E

// gbps / roofline_fraction of a timed launch
void fill_rate(const hsc_handle* h, hsc_stats* s) {
  s->gbps = s->pcg_ms > 0 ? pcg_bytes_per_iteration(h) * s->pcg_iterations / (s->pcg_ms * 1e-3) / 1e9 : 0.0;
  s->roofline_fraction = s->gbps / kHbmPeakGBs;
}

void fill_common_stats(hsc_handle* h, hsc_stats* s) {
  std::memset(s, 0, sizeof(*s));
  s->fem_dofs = h->fn;
  s->abd_dofs = h->qm;
  s->active_fem_dofs = h->a;
  s->rank = h->k;
  s->aca_relative_error = h->aca_err;
  s->adaptive_abd_inverse = h->abd_mode == 2 ? 1 : 0;
  s->neumann_doublings = h->doublings;
  s->fp64_oracle_error = -1.0;
  s->abd_inverse_residual = h->abd_res;
  s->full_residual_limit = h->cfg.full_residual_limit;
  s->precision = h->cfg.precision;
  s->preconditioner = h->cfg.preconditioner;
  s->learned_used = h->learned_used;
  s->fallback = h->fallback_reason != 0;
  s->fallback_reason = h->fallback_reason;
  s->setup_ms = h->setup_ms;
}

// back-substitution + descent + full residual (linear_hsc.cu:395-435)
// Back-substitution of the ABD part and the acceptance reductions (full
// residual, descent) into h->misc[20..22]; kernels only (capturable).
void enqueue_finish(hsc_handle* h, const double* b, double* x, hipStream_t st) {
  HIPCHK(hsc::launch_backsub(h->P.as<double>(), h->U.as<double>(), h->V.as<double>(), h->act_dof.as<int>(), b, x,
                             h->qo, h->qm, h->fo, h->a, h->k, h->small.as<double>(), st));
  hsc::SysDesc S = make_desc(h);
  launch_operator(h, 0, S, x + h->fo, h->yfull.as<double>(), 0.0, st);
  HIPCHK(hsc::launch_full_residual(h->yfull.as<double>(), h->E.as<double>(), h->A.as<double>(),
                                   h->act_node.as<int>(), h->act_dof.as<int>(), h->a, h->qm, b, x, h->fo, h->qo,
                                   h->fn, h->respart.as<double>(), h->misc.as<double>() + 20, st));
}

// Acceptance (linear_hsc.cu:400-415) from out = {|b - Hx|^2, |b|^2, b.x}.
int evaluate_finish(hsc_handle* h, hsc_stats* s, const double out[3]) {
  const double bn = std::sqrt(out[1]);
  s->descent = out[2];
  s->full_relative_residual = bn > 0 ? std::sqrt(out[0]) / bn : 0.0;
  s->accepted = std::isfinite(s->full_relative_residual) && s->full_relative_residual <= h->cfg.full_residual_limit;
  if (h->cfg.check_acceptance) {
    if (!(std::isfinite(out[2]) && (out[2] > 0 || bn == 0))) {
      h->err = "HSC returned a non-descent direction";
      return HSC_ERR_NOT_DESCENT;
    }
    if (!s->accepted) {
      char m[256];
      std::snprintf(m, sizeof m, "HSC full-system residual %g exceeds %g despite Schur residual %g. No fallback.",
                    s->full_relative_residual, h->cfg.full_residual_limit, s->schur_relative_residual);
      h->err = m;
      return HSC_ERR_FULL_RESIDUAL;
    }
  }
  return HSC_OK;
}

int finish_solve(hsc_handle* h, const double* b, double* x, hsc_stats* s, hipStream_t st) {
  enqueue_finish(h, b, x, st);
  double out[3];
  HIPCHK(hipMemcpyAsync(out, h->misc.as<double>() + 20, sizeof(out), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return evaluate_finish(h, s, out);
}

int status_from_out(hsc_handle* h, const hsc::SysOut& o) {
  char m[256];
  if (o.status == HSC_ERR_BREAKDOWN) {
    std::snprintf(m, sizeof m, "HSC Schur PCG breakdown: iteration=%d. No fallback.", o.iterations);
    h->err = m;
  } else if (o.status == HSC_ERR_NOT_CONVERGED) {
    std::snprintf(m, sizeof m, "HSC Schur PCG did not converge: iterations=%d, residual=%g", o.iterations, o.rel);
    h->err = m;
  } else if (o.status != 0) {
    h->err = "persistent PCG: device wait expired";
  }
  return o.status;
}

// Reduced right-hand side and initial residual of h's system (rhs = r).
void enqueue_rhs(hsc_handle* h, const double* b, double* x, hipStream_t st) {
  HIPCHK(hsc::launch_rhs(h->W1.as<double>() + (size_t)h->qm * h->limit, h->V.as<double>(), b, h->qo, h->qm, h->fo,
                         h->fn, h->k, h->a, h->act_node.as<int>(), h->ybuf.as<double>(), h->rhs.as<double>(),
                         h->r.as<double>(), x, st));
}

int solve_impl(hsc_handle* h, const double* b, double* x, hsc_stats* s, hipStream_t st, Clock::time_point t0) {
  if (!h->setup_ok) {
    h->err = "hsc_solve before a successful hsc_setup";
    return HSC_ERR_NOT_SETUP;
  }
  if (!b || !x) throw HscError{HSC_ERR_INVALID_ARGUMENT, "null b or x"};
  hsc_stats tmp;
  if (!s) s = &tmp;
  fill_common_stats(h, s);
  hsc_handle* hs[1] = {h};
  // a reconfiguration after setup that changed the launch re-plans it here
  if (!h->launch_ready) prepare_launch(h, st);
  h->xsol = x + h->fo;
  h->last_xsol = h->xsol;
  const LaunchCfg lc{h->G, h->wgpcu, h->sched};
  enqueue_rhs(h, b, x, st);
  float pms = 0.f;
  std::vector<double> fin;
  auto out = run_pcg(hs, 1, lc, 0, st, &pms, [&] { enqueue_finish(h, b, x, st); }, &fin);
  s->pcg_ms = pms;
  s->pcg_iterations = out[0].iterations;
  fill_rate(h, s);
  s->schur_relative_residual = out[0].rel;
  s->true_residual_checks = out[0].checks;
  s->restarts = out[0].restarts;
  // the PCG's own failure message takes precedence over the acceptance one
  int c2 = evaluate_finish(h, s, fin.data());
  int code = status_from_out(h, out[0]);
  if (code == HSC_OK) code = c2;
  s->status = code;
  s->total_ms = ms_since(t0);
  return code;
}

template <class F>
int guarded(hsc_handle* h, F&& f) {
  try {
    return f();
  } catch (const HscError& e) {
    if (h) h->err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    if (h) h->err = e.what();
    return HSC_ERR_HIP;
  }
}

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

}  // namespace

extern "C" {

int32_t hsc_abi_version(void) { return HSC_ABI_VERSION; }

void hsc_config_default(hsc_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->max_rank = 128;
  c->neumann_order = 5;
  c->max_iter = 2000;
  c->adaptive_abd_inverse = 1;
  c->abd_inverse_tol = 1e-5;
  c->full_residual_limit = 0.1;
  c->omega = 0.7;
  c->tol_rate = 1e-3;
  c->abd_inverse = HSC_ABD_NEUMANN;  // the plugin's polynomial (+ doubling: adaptive_abd_inverse)
  c->precision = HSC_PRECISION_FP64;
  c->residual_replace_every = 25;
  c->true_residual_every_iteration = 0;
  c->aca_rtol = 0.0;
  c->lanes_per_row = 0;
  c->wg_per_cu = 0;
  c->schedule = 0;
  c->preconditioner = HSC_PRECOND_NEUMANN;
  c->check_acceptance = 1;
  c->poly_interval_ratio = 100.0;
  c->pass_sync = 2;
  c->pipelined = 0;
  c->operator_mapping = 0;
  c->reserved0 = 0;
}

int32_t hsc_create(const hsc_config* cfg, int32_t device, hsc_handle** out) {
  if (!out) return HSC_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  hsc_config c;
  if (cfg) c = *cfg;
  else hsc_config_default(&c);
  std::string why;
  int v = validate_config(c, why);
  if (v != HSC_OK) return v;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return HSC_ERR_HIP;
  hsc_handle* h = new hsc_handle();
  h->cfg = c;
  h->device = device;
  *out = h;
  return HSC_OK;
}

void hsc_destroy(hsc_handle* h) {
  if (!h) return;
  DeviceGuard g(h->device);
  delete h;
}

const char* hsc_last_error(const hsc_handle* h) { return h ? h->err.c_str() : "null handle"; }

int32_t hsc_set_config(hsc_handle* h, const hsc_config* cfg) {
  if (!h || !cfg) return HSC_ERR_INVALID_ARGUMENT;
  std::string why;
  int v = validate_config(*cfg, why);
  if (v != HSC_OK) {
    h->err = why;
    return v;
  }
  const bool refactor = cfg->max_rank != h->cfg.max_rank || cfg->abd_inverse != h->cfg.abd_inverse ||
                        cfg->neumann_order != h->cfg.neumann_order || cfg->omega != h->cfg.omega ||
                        cfg->aca_rtol != h->cfg.aca_rtol || cfg->adaptive_abd_inverse != h->cfg.adaptive_abd_inverse ||
                        cfg->abd_inverse_tol != h->cfg.abd_inverse_tol ||
                        cfg->preconditioner != h->cfg.preconditioner ||
                        cfg->poly_interval_ratio != h->cfg.poly_interval_ratio;
  // the launch shape and its scratch depend on these (re-planned by the next solve)
  const bool relaunch = cfg->precision != h->cfg.precision || cfg->lanes_per_row != h->cfg.lanes_per_row ||
                        cfg->wg_per_cu != h->cfg.wg_per_cu || cfg->schedule != h->cfg.schedule ||
                        cfg->pass_sync != h->cfg.pass_sync || cfg->pipelined != h->cfg.pipelined;
  h->cfg = *cfg;
  ++h->gen;  // a captured batch of this handle is stale now
  if (refactor) h->setup_ok = false;
  if (relaunch) h->launch_ready = false;
  return HSC_OK;
}

int32_t hsc_setup(hsc_handle* h, const hsc_bcoo3_view* A, int32_t qo, int32_t qm, int32_t fo, int32_t fn,
                  hsc_stream stream) {
  if (!h) return HSC_ERR_INVALID_ARGUMENT;
  DeviceGuard g(h->device);
  return guarded(h, [&] {
    do_setup(h, A, qo, qm, fo, fn, (hipStream_t)stream);
    return HSC_OK;
  });
}

int32_t hsc_solve(hsc_handle* h, const double* b, double* x, hsc_stats* stats, hsc_stream stream) {
  if (!h) return HSC_ERR_INVALID_ARGUMENT;
  DeviceGuard g(h->device);
  auto t0 = Clock::now();
  int code = guarded(h, [&] { return solve_impl(h, b, x, stats, (hipStream_t)stream, t0); });
  if (stats) {
    stats->status = code;
    stats->setup_ms = 0.0;
    stats->total_ms = ms_since(t0);
  }
  return code;
}

int32_t hsc_do_solve(hsc_handle* h, const hsc_bcoo3_view* A, int32_t qo, int32_t qm, int32_t fo, int32_t fn,
                     const double* b, double* x, hsc_stats* stats, hsc_stream stream) {
  if (!h) return HSC_ERR_INVALID_ARGUMENT;
  DeviceGuard g(h->device);
  auto t0 = Clock::now();
  int code = guarded(h, [&] {
    do_setup(h, A, qo, qm, fo, fn, (hipStream_t)stream);
    return HSC_OK;
  });
  if (code == HSC_OK) code = guarded(h, [&] { return solve_impl(h, b, x, stats, (hipStream_t)stream, t0); });
  if (stats) {
    if (code != HSC_OK && !h->setup_ok) fill_common_stats(h, stats);
    stats->status = code;
    stats->setup_ms = h->setup_ok ? h->setup_ms : ms_since(t0);
    stats->total_ms = ms_since(t0);
  }
  return code;
}

int32_t hsc_solve_batched(hsc_handle* const* handles, int32_t count, const double* const* b, double* const* x,
                          hsc_stats* stats, hsc_stream stream) {
  if (!handles || count <= 0 || !b || !x) return HSC_ERR_INVALID_ARGUMENT;
  if (count > hsc::kMaxSys) return HSC_ERR_CAPACITY;
  hsc_handle* h0 = handles[0];
  if (!h0) return HSC_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < count; ++i) {
    if (!handles[i] || !handles[i]->setup_ok) return HSC_ERR_NOT_SETUP;
    if (handles[i]->device != h0->device || handles[i]->cfg.precision != h0->cfg.precision ||
        handles[i]->cfg.neumann_order != h0->cfg.neumann_order || handles[i]->cfg.omega != h0->cfg.omega)
      return HSC_ERR_INVALID_ARGUMENT;
  }
  DeviceGuard g(h0->device);
  hipStream_t st = (hipStream_t)stream;
  auto t0 = Clock::now();
  return guarded(h0, [&] {
    for (int i = 0; i < count; ++i) handles[i]->xsol = x[i] + handles[i]->fo;
    LaunchCfg lc{};
    if (!tune_lookup(handles, count, &lc)) {
      // the timing runs start from the real right-hand sides
      for (int i = 0; i < count; ++i) enqueue_rhs(handles[i], b[i], x[i], st);
      lc = tune_now(handles, count, st);
    }
    for (int i = 0; i < count; ++i) {
      hsc_handle* h = handles[i];
      enqueue_rhs(h, b[i], x[i], st);
    }
    float pms = 0.f;
    std::vector<double> fin;
    auto out = run_pcg(
        handles, count, lc, 0, st, &pms,
        [&] {
          for (int i = 0; i < count; ++i) enqueue_finish(handles[i], b[i], x[i], st);
        },
        &fin);
    int first = HSC_OK;
    for (int i = 0; i < count; ++i) {
      hsc_handle* h = handles[i];
      hsc_stats tmp;
      hsc_stats* s = stats ? &stats[i] : &tmp;
      fill_common_stats(h, s);
      s->setup_ms = 0.0;
      s->pcg_ms = pms;
      s->pcg_iterations = out[i].iterations;
      s->schur_relative_residual = out[i].rel;
      s->true_residual_checks = out[i].checks;
      s->restarts = out[i].restarts;
      int c2 = evaluate_finish(h, s, fin.data() + 3 * i);
      int code = status_from_out(h, out[i]);
      if (code == HSC_OK) code = c2;
      s->status = code;
      s->total_ms = ms_since(t0);
      if (first == HSC_OK) first = code;
    }
    return first;
  });
}

}  // extern "C"

// A batched solve captured once into a hipGraph (right-hand sides, state
// reset, the persistent PCG launch, back-substitution and acceptance
// reductions of every system) and replayed for fixed b/x buffers.
struct hsc_batch {
  std::vector<hsc_handle*> hs;
  std::vector<const double*> b;
  std::vector<double*> x;
  PcgScratch sc;  // own launch scratch: eager solves on the handles do not touch the captured tables
  PcgPlan plan;
  std::vector<unsigned long long> gens;  // each handle's setup generation at capture
  LaunchCfg lc{8, 1, 1};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int device = 0;
  int cooperative_node = 0;
  ~hsc_batch() {
    DeviceGuard g(device);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

namespace {

int validate_batch(hsc_handle* const* handles, int32_t count, const double* const* b, double* const* x) {
  if (!handles || count <= 0 || !b || !x) return HSC_ERR_INVALID_ARGUMENT;
  if (count > hsc::kMaxSys) return HSC_ERR_CAPACITY;
  hsc_handle* h0 = handles[0];
  if (!h0) return HSC_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < count; ++i) {
    if (!handles[i] || !handles[i]->setup_ok) return HSC_ERR_NOT_SETUP;
    if (!b[i] || !x[i]) return HSC_ERR_INVALID_ARGUMENT;
    if (handles[i]->device != h0->device || handles[i]->cfg.precision != h0->cfg.precision ||
        handles[i]->cfg.neumann_order != h0->cfg.neumann_order || handles[i]->cfg.omega != h0->cfg.omega ||
        handles[i]->cfg.preconditioner != h0->cfg.preconditioner)
      return HSC_ERR_INVALID_ARGUMENT;
    for (int j = 0; j < i; ++j)
      if (handles[j] == handles[i]) return HSC_ERR_INVALID_ARGUMENT;  // one system per handle
  }
  return HSC_OK;
}

// HSC_ERR_STALE_BATCH when a handle was set up again or reconfigured after
// the capture (its buffers, rank and tile ranges may no longer be the ones
// the graph holds), HSC_ERR_NOT_SETUP when its last setup failed.
int batch_stale(hsc_batch* bt) {
  for (size_t i = 0; i < bt->hs.size(); ++i) {
    hsc_handle* h = bt->hs[i];
    if (h->gen != bt->gens[i]) {
      bt->hs[0]->err = "captured batch: system " + std::to_string(i) + " was set up or reconfigured after the capture";
      return HSC_ERR_STALE_BATCH;
    }
    if (!h->setup_ok) return HSC_ERR_NOT_SETUP;
  }
  return HSC_OK;
}


}  // namespace

extern "C" {

int32_t hsc_batch_create(hsc_handle* const* handles, int32_t count, const double* const* b, double* const* x,
                         hsc_stream stream, hsc_batch** out) {
  if (!out) return HSC_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  const int v = validate_batch(handles, count, b, x);
  if (v != HSC_OK) return v;
  hsc_handle* h0 = handles[0];
  DeviceGuard g(h0->device);
  hipStream_t st = (hipStream_t)stream;
  return guarded(h0, [&] {
    std::unique_ptr<hsc_batch> bt(new hsc_batch());
    bt->device = h0->device;
    bt->hs.assign(handles, handles + count);
    bt->b.assign(b, b + count);
    bt->x.assign(x, x + count);
    for (int i = 0; i < count; ++i) handles[i]->xsol = x[i] + handles[i]->fo;
    if (!tune_lookup(handles, count, &bt->lc)) {
      for (int i = 0; i < count; ++i) enqueue_rhs(handles[i], b[i], x[i], st);
      bt->lc = tune_now(handles, count, st);
    }
    for (int i = 0; i < count; ++i) bt->gens.push_back(handles[i]->gen);
    bt->plan = plan_pcg(handles, count, bt->lc, 0, st, bt->sc);
    HIPCHK(hipStreamSynchronize(st));
    hipStream_t cap = nullptr;
    HIPCHK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    hipError_t ce = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    std::string what;
    if (ce == hipSuccess) {
      try {
        for (int i = 0; i < count; ++i) enqueue_rhs(handles[i], b[i], x[i], cap);
        reset_pcg(handles, count, bt->sc, cap);
        HIPCHK(hsc::launch_pcg_plain(bt->plan.G, bt->plan.mixed, bt->plan.res, bt->plan.pipe, bt->plan.P,
                                     bt->plan.shmem, cap));
        for (int i = 0; i < count; ++i) enqueue_finish(handles[i], b[i], x[i], cap);
      } catch (const HscError& e) {
        what = e.msg;
      }
      hipGraph_t graph = nullptr;
      const hipError_t ee = hipStreamEndCapture(cap, &graph);
      bt->graph = graph;
      if (what.empty() && ee != hipSuccess) what = std::string("graph capture: ") + hipGetErrorString(ee);
    } else {
      what = std::string("graph capture: ") + hipGetErrorString(ce);
    }
    (void)hipStreamDestroy(cap);
    if (!what.empty()) throw HscError{HSC_ERR_HIP, what};
    // mark the persistent PCG node cooperative (co-residency of its grid)
    const void* fn = hsc::pcg_kernel_fn(bt->plan.G, bt->plan.mixed, bt->plan.res, bt->plan.pipe);
    size_t nn = 0;
    HIPCHK(hipGraphGetNodes(bt->graph, nullptr, &nn));
    std::vector<hipGraphNode_t> nodes(nn);
    HIPCHK(hipGraphGetNodes(bt->graph, nodes.data(), &nn));
    hipGraphNode_t pcg_node = nullptr;
    for (auto nd : nodes) {
      hipGraphNodeType t;
      if (hipGraphNodeGetType(nd, &t) != hipSuccess || t != hipGraphNodeTypeKernel) continue;
      hipKernelNodeParams kp{};
      if (hipGraphKernelNodeGetParams(nd, &kp) == hipSuccess && kp.func == fn) pcg_node = nd;
    }
    if (!pcg_node) throw HscError{HSC_ERR_HIP, "graph capture: persistent PCG node not found"};
    hipKernelNodeAttrValue av{};
    av.cooperative = 1;
    bt->cooperative_node = hipGraphKernelNodeSetAttribute(pcg_node, hipKernelNodeAttributeCooperative, &av) ==
                           hipSuccess;
    (void)hipGetLastError();
    hipError_t ie = hipGraphInstantiate(&bt->exec, bt->graph, nullptr, nullptr, 0);
    if (ie != hipSuccess && bt->cooperative_node) {
      (void)hipGetLastError();
      av.cooperative = 0;
      (void)hipGraphKernelNodeSetAttribute(pcg_node, hipKernelNodeAttributeCooperative, &av);
      bt->cooperative_node = 0;
      ie = hipGraphInstantiate(&bt->exec, bt->graph, nullptr, nullptr, 0);
    }
    if (ie != hipSuccess) throw HscError{HSC_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ie)};
    *out = bt.release();
    return HSC_OK;
  });
}

int32_t hsc_batch_launch(hsc_batch* bt, hsc_stream stream) {
  if (!bt || !bt->exec) return HSC_ERR_INVALID_ARGUMENT;
  const int stale = batch_stale(bt);
  if (stale != HSC_OK) return stale;
  DeviceGuard g(bt->device);
  const hipError_t e = hipGraphLaunch(bt->exec, (hipStream_t)stream);
  if (e != hipSuccess) {
    bt->hs[0]->err = std::string("graph launch: ") + hipGetErrorString(e);
    return HSC_ERR_HIP;
  }
  return HSC_OK;
}

int32_t hsc_batch_stats(hsc_batch* bt, hsc_stats* stats, hsc_stream stream) {
  if (!bt || !stats) return HSC_ERR_INVALID_ARGUMENT;
  const int stale = batch_stale(bt);
  if (stale != HSC_OK) return stale;
  DeviceGuard g(bt->device);
  hipStream_t st = (hipStream_t)stream;
  const int n = (int)bt->hs.size();
  hsc_handle* h0 = bt->hs[0];
  return guarded(h0, [&] {
    std::vector<double> res;
    std::vector<hsc::SysOut> out = collect_pcg(bt->sc, n, st, bt->hs.data(), &res);
    int first = HSC_OK;
    for (int i = 0; i < n; ++i) {
      hsc_handle* h = bt->hs[i];
      hsc_stats* s = &stats[i];
      fill_common_stats(h, s);
      s->pcg_iterations = out[i].iterations;
      s->schur_relative_residual = out[i].rel;
      s->true_residual_checks = out[i].checks;
      s->restarts = out[i].restarts;
      const int c2 = evaluate_finish(h, s, res.data() + 3 * i);
      int code = status_from_out(h, out[i]);
      if (code == HSC_OK) code = c2;
      s->status = code;
      if (first == HSC_OK) first = code;
    }
    return first;
  });
}

int32_t hsc_batch_info(const hsc_batch* bt, int32_t* lanes_per_row, int32_t* wg_per_cu, int32_t* workgroups,
                       int32_t* cooperative_node, int32_t* schedule) {
  if (!bt) return HSC_ERR_INVALID_ARGUMENT;
  if (lanes_per_row) *lanes_per_row = bt->lc.G;
  if (wg_per_cu) *wg_per_cu = bt->lc.wgpcu;
  if (workgroups) *workgroups = bt->plan.P.nwg;
  if (cooperative_node) *cooperative_node = bt->cooperative_node;
  if (schedule) *schedule = bt->lc.sched;
  return HSC_OK;
}

void hsc_batch_destroy(hsc_batch* bt) { delete bt; }

const char* hsc_preconditioner_name(int32_t p) {
  switch (p) {
    case HSC_PRECOND_NEUMANN: return "neumann";
    case HSC_PRECOND_LEARNED: return "learned";
    case HSC_PRECOND_CHEBYSHEV: return "chebyshev";
    default: return "unknown";
  }
}

const char* hsc_precision_name(int32_t p) {
  switch (p) {
    case HSC_PRECISION_FP64: return "fp64";
    case HSC_PRECISION_MIXED: return "mixed";
    default: return "unknown";
  }
}

int32_t hsc_append_jsonl(const hsc_stats* s, const char* path, int32_t frame, int32_t newton) {
  if (!s || !path) return HSC_ERR_INVALID_ARGUMENT;
  FILE* f = std::fopen(path, "a");
  if (!f) return HSC_ERR_INVALID_ARGUMENT;
  // the reference's fields in the order of linear_hsc.cu:421-431, then the additions
  char oracle[40];
  if (s->fp64_oracle_error >= 0 && std::isfinite(s->fp64_oracle_error))
    std::snprintf(oracle, sizeof oracle, "%.17g", s->fp64_oracle_error);
  else
    std::snprintf(oracle, sizeof oracle, "null");
  std::fprintf(f,
               "{\"frame\":%d,\"newton\":%d,\"fem_dofs\":%d,\"abd_dofs\":%d,\"active_fem_dofs\":%d,\"rank\":%d,"
               "\"aca_relative_error\":%.17g,\"pcg_iterations\":%d,\"schur_relative_residual\":%.17g,"
               "\"full_relative_residual\":%.17g,\"setup_ms\":%.6g,\"total_ms\":%.6g,\"active_mesh_contacts\":%d,"
               "\"active_ground_contacts\":%d,\"adaptive_abd_inverse\":%s,\"neumann_doublings\":%d,"
               "\"abd_inverse_residual\":%.17g,\"full_residual_limit\":%.17g,\"accepted\":%s,"
               "\"solver\":\"linear_hsc\",\"fallback\":false,\"precision\":\"%s\",\"preconditioner\":\"%s\","
               "\"learned_used\":%s,\"learned_fallback\":%s,\"fallback_reason\":%d,\"status\":%d,\"pcg_ms\":%.6g,"
               "\"gbps\":%.6g,\"roofline_fraction\":%.6g,\"fp64_oracle_error\":%s}\n",
               frame, newton, s->fem_dofs, s->abd_dofs, s->active_fem_dofs, s->rank, s->aca_relative_error,
               s->pcg_iterations, s->schur_relative_residual, s->full_relative_residual, s->setup_ms, s->total_ms,
               s->active_mesh_contacts, s->active_ground_contacts, s->adaptive_abd_inverse ? "true" : "false",
               s->neumann_doublings, s->abd_inverse_residual, s->full_residual_limit, s->accepted ? "true" : "false",
               hsc_precision_name(s->precision), hsc_preconditioner_name(s->preconditioner),
               s->learned_used ? "true" : "false", s->fallback ? "true" : "false", s->fallback_reason, s->status,
               s->pcg_ms, s->gbps, s->roofline_fraction, oracle);
  std::fclose(f);
  return HSC_OK;
}

int32_t hsc_tune_file(const char* path) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tune_path = path ? path : "";
  g_tune_path_set = true;
  g_tune_loaded = false;  // (re)load from the new file; entries in memory are kept
  load_tune_locked();
  return HSC_OK;
}

int32_t hsc_tune(hsc_handle* h, hsc_stream stream) {
  if (!h) return HSC_ERR_INVALID_ARGUMENT;
  if (!h->setup_ok) return HSC_ERR_NOT_SETUP;
  DeviceGuard g(h->device);
  return guarded(h, [&] {
    prepare_launch(h, (hipStream_t)stream, true);
    return HSC_OK;
  });
}

int64_t hsc_alloc_count(void) { return g_alloc_count.load(); }

int32_t hsc_apply_D(hsc_handle* h, const double* x, double* y, hsc_stream stream) {
  if (!h || !x || !y) return HSC_ERR_INVALID_ARGUMENT;
  if (!h->setup_ok) return HSC_ERR_NOT_SETUP;
  DeviceGuard g(h->device);
  return guarded(h, [&] {
    hsc::SysDesc S = make_desc(h);
    hipStream_t st = (hipStream_t)stream;
    launch_operator(h, 0, S, x, y, 0.0, st);
    HIPCHK(hipStreamSynchronize(st));
    return HSC_OK;
  });
}

int32_t hsc_apply_schur(hsc_handle* h, const double* x, double* y, hsc_stream stream) {
  if (!h || !x || !y) return HSC_ERR_INVALID_ARGUMENT;
  if (!h->setup_ok) return HSC_ERR_NOT_SETUP;
  DeviceGuard g(h->device);
  return guarded(h, [&] {
    hsc::SysDesc S = make_desc(h);
    hipStream_t st = (hipStream_t)stream;
    launch_operator(h, 0, S, x, y, 0.0, st);
    HIPCHK(hsc::launch_op_lowrank64(S, x, y, h->ybuf.as<double>(), st));
    HIPCHK(hipStreamSynchronize(st));
    return HSC_OK;
  });
}

int32_t hsc_apply_precond(hsc_handle* h, const double* r, double* z, hsc_stream stream) {
  if (!h || !r || !z) return HSC_ERR_INVALID_ARGUMENT;
  if (!h->setup_ok) return HSC_ERR_NOT_SETUP;
  DeviceGuard g(h->device);
  return guarded(h, [&] {
    hsc::SysDesc S = make_desc(h);
    hipStream_t st = (hipStream_t)stream;
    const hsc::StripDesc T = strip_desc(h);
    HIPCHK(hsc::launch_op_precond64(streaming_ops(h, 1) ? &T : nullptr, 16, S, h->cfg.neumann_order, h->omegas, r,
                                    z, h->z1.as<double>(), st));
    HIPCHK(hipStreamSynchronize(st));
    return HSC_OK;
  });
}

int32_t hsc_apply_pass(hsc_handle* h, int32_t precision, double w, const void* x, const double* r, void* y,
                       hsc_stream stream) {
  if (!h || !x || !r || !y || (precision != HSC_PRECISION_FP64 && precision != HSC_PRECISION_MIXED) ||
      !std::isfinite(w))
    return HSC_ERR_INVALID_ARGUMENT;
  if (!h->setup_ok) return HSC_ERR_NOT_SETUP;
  DeviceGuard g(h->device);
  return guarded(h, [&] {
    hsc::SysDesc S = make_desc(h);
    S.r = const_cast<double*>(r);
    hipStream_t st = (hipStream_t)stream;
    launch_operator(h, precision == HSC_PRECISION_FP64 ? 1 : 2, S, x, y, w, st);
    HIPCHK(hipStreamSynchronize(st));
    return HSC_OK;
  });
}

int32_t hsc_fem_hessian(const int32_t* tets, const double* x, const double* Dm_inv, const double* mu,
                        const double* lambda, const double* volumes, double dt, int64_t n_tets, int32_t project_spd,
                        int32_t* out_rows, int32_t* out_cols, double* out_values, hsc_stream stream) {
  if (n_tets < 0 || (n_tets > 0 && (!tets || !x || !Dm_inv || !mu || !lambda || !volumes || !out_rows ||
                                    !out_cols || !out_values)))
    return HSC_ERR_INVALID_ARGUMENT;
  if (n_tets > (1LL << 27)) return HSC_ERR_CAPACITY;
  hsc::FemHessianParams P{tets, x, Dm_inv, mu, lambda, volumes, dt * dt, n_tets, project_spd ? 1 : 0,
                          out_rows, out_cols, out_values};
  hipError_t e = hsc::launch_fem_hessian(P, (hipStream_t)stream);
  if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
  if (e != hipSuccess) std::fprintf(stderr, "hsc_fem_hessian: %s\n", hipGetErrorString(e));
  return e == hipSuccess ? HSC_OK : HSC_ERR_HIP;
}

int32_t hsc_abd_fem_coupling(const int32_t* abd_vertex, const int32_t* fem_vertex, const double* hess,
                             int64_t count, const int32_t* v2b, const double* xbar, int32_t n_abd_vertices,
                             const int32_t* body_fixed, int32_t n_bodies, const int32_t* vertex_fixed,
                             int32_t n_fem_vertices, int32_t abd_block_offset, int32_t fem_block_offset,
                             int32_t* out_rows, int32_t* out_cols, double* out_values, hsc_stream stream) {
  if (count < 0 || n_abd_vertices < 0 || n_bodies < 0 || n_fem_vertices < 0 || abd_block_offset < 0 ||
      fem_block_offset < 0)
    return HSC_ERR_INVALID_ARGUMENT;
  if (count == 0) return HSC_OK;
  if (!abd_vertex || !fem_vertex || !hess || !v2b || !xbar || !body_fixed || !vertex_fixed || !out_rows ||
      !out_cols || !out_values)
    return HSC_ERR_INVALID_ARGUMENT;
  if (count > (1LL << 28) || (long long)abd_block_offset + 4LL * n_bodies > INT32_MAX ||
      (long long)fem_block_offset + n_fem_vertices > INT32_MAX)
    return HSC_ERR_CAPACITY;
  hipStream_t st = (hipStream_t)stream;
  // the index-error word: one device int and one pinned host int per device
  // (the device of `stream`), allocated on the first call and kept; calls are
  // serialised by the lock
  static std::mutex mu;
  static std::map<int, std::pair<int*, int*>> flags;
  std::lock_guard<std::mutex> lock(mu);
  hipDevice_t dev = 0;
  hipError_t e = hipStreamGetDevice(st, &dev);
  if (e != hipSuccess) return HSC_ERR_HIP;
  DeviceGuard g(dev);
  auto& f = flags[dev];
  if (!f.first) {
    e = hipMalloc((void**)&f.first, sizeof(int));
    if (e == hipSuccess) hsc::note_alloc();
    if (e == hipSuccess) e = hipHostMalloc((void**)&f.second, sizeof(int), hipHostMallocDefault);
    if (e == hipSuccess) hsc::note_alloc();
    if (e != hipSuccess) {
      if (f.first) (void)hipFree(f.first);
      f = {nullptr, nullptr};
      return HSC_ERR_HIP;
    }
  }
  int* err = f.first;
  e = hipMemsetAsync(err, 0, sizeof(int), st);
  if (e == hipSuccess) {
    hsc::AbdFemCouplingParams P{abd_vertex,     fem_vertex,     hess,     v2b,      xbar,
                                body_fixed,     vertex_fixed,   count,    n_abd_vertices,
                                n_fem_vertices, n_bodies,       abd_block_offset,
                                fem_block_offset, err,          out_rows, out_cols, out_values};
    e = hsc::launch_abd_fem_coupling(P, st);
  }
  if (e == hipSuccess) e = hipMemcpyAsync(f.second, err, sizeof(int), hipMemcpyDeviceToHost, st);
  const hipError_t se = hipStreamSynchronize(st);
  if (e == hipSuccess) e = se;
  if (e != hipSuccess) {
    std::fprintf(stderr, "hsc_abd_fem_coupling: %s\n", hipGetErrorString(e));
    return HSC_ERR_HIP;
  }
  return *f.second ? HSC_ERR_INVALID_ARGUMENT : HSC_OK;
}

int32_t hsc_assemble_bcoo(const int32_t* rows, const int32_t* cols, const double* values, int64_t count,
                          int32_t fold_lower, int32_t* out_rows, int32_t* out_cols, double* out_values,
                          int64_t* out_count, hsc_stream stream) {
  if (!out_count || count < 0 || (count > 0 && (!rows || !cols || !values || !out_rows || !out_cols || !out_values)))
    return HSC_ERR_INVALID_ARGUMENT;
  if (count >= (1LL << 31)) return HSC_ERR_CAPACITY;
  long long n = 0;
  const hipError_t e = hsc::assemble_bcoo(rows, cols, values, count, fold_lower ? 1 : 0, out_rows, out_cols,
                                          out_values, &n, (hipStream_t)stream);
  if (e != hipSuccess) std::fprintf(stderr, "hsc_assemble_bcoo: %s\n", hipGetErrorString(e));
  *out_count = n;
  return e == hipSuccess ? HSC_OK : HSC_ERR_HIP;
}

int32_t hsc_time_kernel(hsc_handle* h, int32_t kernel, int32_t lanes_per_row, int32_t repeats, hsc_stream stream,
                        double* ms_per_launch) {
  if (!h || !ms_per_launch || kernel < 0 || kernel > 5 || repeats < 1) return HSC_ERR_INVALID_ARGUMENT;
  if (kernel < 3 && !(lanes_per_row == 4 || lanes_per_row == 8 || lanes_per_row == 16))
    return HSC_ERR_INVALID_ARGUMENT;
  if (kernel >= 3 && (lanes_per_row < 0 || lanes_per_row > 64)) return HSC_ERR_INVALID_ARGUMENT;
  if (!h->setup_ok) return HSC_ERR_NOT_SETUP;
  DeviceGuard g(h->device);
  return guarded(h, [&] {
    hipStream_t st = (hipStream_t)stream;
    hsc::SysDesc S = make_desc(h);
    // the handle's own vectors as operands (zeroed; the arithmetic does not
    // depend on the values)
    const size_t vb = sizeof(double) * (size_t)h->fn;
    const void* zp[] = {h->p0.p, h->p1.p, h->r.p};
    const size_t zb[] = {vb, vb, vb};
    HIPCHK(hsc::launch_zero_multi(zp, zb, 3, st));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    hipError_t le = hipSuccess;
    (void)hipEventRecord(e0, st);
    hsc::StripDesc T{};
    if (kernel >= 3) {
      T = strip_desc(h);
      if (lanes_per_row > 0) T.nwave = (T.nch + lanes_per_row - 1) / lanes_per_row;
    }
    for (int i = 0; i < repeats && le == hipSuccess; ++i)
      le = kernel < 3 ? hsc::launch_row_op(kernel, lanes_per_row, S, h->p0.p, h->p1.p, 0.5, st)
                      : hsc::launch_strip_op(kernel - 3, S, T, h->p0.p, h->p1.p, 0.5, st);
    (void)hipEventRecord(e1, st);
    hipError_t se = hipEventSynchronize(e1);
    float ms = 0.f;
    if (le == hipSuccess && se == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHK(le);
    HIPCHK(se);
    *ms_per_launch = ms / repeats;
    return HSC_OK;
  });
}

int32_t hsc_get_setup_info(const hsc_handle* h, hsc_setup_info* o) {
  if (!h || !o) return HSC_ERR_INVALID_ARGUMENT;
  std::memset(o, 0, sizeof(*o));
  o->fem_dofs = h->fn;
  o->abd_dofs = h->qm;
  o->fem_nodes = h->nb;
  o->nnzb = h->nnzb;
  o->active_fem_dofs = h->a;
  o->rank = h->k;
  o->lanes_per_row = h->G;
  o->workgroups = h->wgpcu;
  o->schedule = h->sched;
  o->pass_sync = h->pass_sync_used ? 1 : 0;
  o->pipelined = h->pipe_used ? 1 : 0;
  o->pass_pairs = h->pairs_used ? 1 : 0;
  o->streaming_operators = streaming_ops(h, 0) ? 1 : 0;
  o->row_tiles = h->ntiles;
  o->lambda_max = h->lambda_max;
  o->neumann_doublings = h->doublings;
  o->abd_block_diagonal = h->blockdiag ? 1 : 0;
  o->aca_relative_error = h->aca_err;
  o->abd_inverse_residual = h->abd_res;
  o->poly_ratio = h->poly_ratio;
  for (int c = 0; c < 3; ++c) o->learned_out[c] = h->learned_out[c];
  o->learned_used = h->learned_used;
  o->fallback_reason = h->fallback_reason;
  return h->setup_ok ? HSC_OK : HSC_ERR_NOT_SETUP;
}

int32_t hsc_get_factors(hsc_handle* h, double* U, double* V, double* C, double* P, int32_t* active,
                        double* aca_history) {
  if (!h) return HSC_ERR_INVALID_ARGUMENT;
  if (!h->setup_ok) return HSC_ERR_NOT_SETUP;
  DeviceGuard g(h->device);
  return guarded(h, [&] {
    if (U && h->k) HIPCHK(hipMemcpy(U, h->U.p, sizeof(double) * h->qm * h->k, hipMemcpyDeviceToHost));
    if (V && h->k) HIPCHK(hipMemcpy(V, h->V.p, sizeof(double) * (size_t)h->a * h->k, hipMemcpyDeviceToHost));
    if (C && h->k) HIPCHK(hipMemcpy(C, h->C.p, sizeof(double) * h->k * h->k, hipMemcpyDeviceToHost));
    if (P) HIPCHK(hipMemcpy(P, h->P.p, sizeof(double) * h->qm * h->qm, hipMemcpyDeviceToHost));
    if (active && h->a) HIPCHK(hipMemcpy(active, h->act_dof.p, sizeof(int) * h->a, hipMemcpyDeviceToHost));
    if (aca_history) HIPCHK(hipMemcpy(aca_history, h->hist.p, sizeof(double) * (h->k + 1), hipMemcpyDeviceToHost));
    return HSC_OK;
  });
}

int32_t hsc_set_learned_weights(hsc_handle* h, const void* w, int64_t bytes) {
  if (!h || (!w && bytes)) return HSC_ERR_INVALID_ARGUMENT;
  if (bytes != 0 && bytes != (int64_t)sizeof(float) * hsc::MlpLayout::count) {
    h->err = "learned weights: expected hsc_mlp_param_count() float32 values";
    return HSC_ERR_INVALID_ARGUMENT;
  }
  DeviceGuard g(h->device);
  return guarded(h, [&] {
    HIPCHK(h->learned.ensure((size_t)std::max<int64_t>(bytes, 4)));
    if (bytes) HIPCHK(hipMemcpy(h->learned.p, w, (size_t)bytes, hipMemcpyDefault));
    h->learned_bytes = bytes;
    ++h->learned_version;
    return HSC_OK;
  });
}

int32_t hsc_get_node_graph(hsc_handle* h, int32_t* nbr_ptr, int32_t* nbr, int64_t* n_edges) {
  if (!h || !n_edges) return HSC_ERR_INVALID_ARGUMENT;
  if (!h->setup_ok) return HSC_ERR_NOT_SETUP;
  DeviceGuard g(h->device);
  return guarded(h, [&] {
    *n_edges = (int64_t)h->nnzb - h->nb;
    if (nbr_ptr || nbr) {
      ensure_graph(h, nullptr);
      if (nbr_ptr) HIPCHK(hipMemcpy(nbr_ptr, h->nbr_ptr.p, sizeof(int) * (h->nb + 1), hipMemcpyDefault));
      if (nbr && *n_edges) HIPCHK(hipMemcpy(nbr, h->nbr.p, sizeof(int) * (size_t)*n_edges, hipMemcpyDefault));
    }
    return HSC_OK;
  });
}

int32_t hsc_get_learned_features(hsc_handle* h, uint16_t* X) {
  if (!h || !X) return HSC_ERR_INVALID_ARGUMENT;
  if (!h->setup_ok) return HSC_ERR_NOT_SETUP;
  DeviceGuard g(h->device);
  return guarded(h, [&] {
    ensure_graph(h, nullptr);
    HIPCHK(h->xfeat.ensure(sizeof(hsc::bf16_t) * 16 * h->nb));
    enqueue_learned_features(h, nullptr);
    HIPCHK(hipMemcpy(X, h->xfeat.p, sizeof(hsc::bf16_t) * 16 * h->nb, hipMemcpyDefault));
    return HSC_OK;
  });
}

}  // extern "C"
