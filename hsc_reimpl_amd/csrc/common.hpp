// Shared host/device definitions of the MI355X HSC solver.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace hsc {

// Counts one device or pinned-host allocation made by the library
// (hsc_alloc_count reports the total).
void note_alloc();

constexpr int kBlock = 512;        // threads per workgroup (8 wave64)
constexpr int kPipeRows = 128;     // rows (= nodes) of a resident tile in the pipelined PCG loop (G = 4)
constexpr int kNodeTileMax = 512;  // nodes per elementwise tile (one node per thread; runtime 128 or 512)
constexpr int kLrTile = 128;       // low-rank rows (W rows) per tile
constexpr int kMaxSys = 128;       // systems per persistent launch
constexpr int kMaxRankLds = 256;   // rank limit of the low-rank apply (lanes hold columns l, l+64, l+128, l+192)
constexpr int kBarrierWords = 1024;  // grid-barrier counters (see device_util.hpp)
constexpr int kMaxPolySteps = 32;    // per-step weights of the polynomial preconditioner
constexpr int kNbrStride = 64;       // row-tile neighbour list: [count, up to 63 tiles]; count -1 = every tile
constexpr int kNbrMaxTiles = 1024;   // row tiles per system for which neighbour lists are built

// Pass pairs (FP32 resident rows, 4 lanes per row, 128-row tiles): a row
// tile computes pass j on its rows and on its halo rows (the rows of other
// tiles that its rows reference), keeps both in LDS and computes pass j+1 on
// its rows from there, so it reads its neighbours' values once per two
// passes (as tagged words, see the kernel).  The tile's node set lists the nodes those passes read: its own
// rows (slots 0..127), its halo rows (128..127+nh), then the other columns
// of the halo rows; each is gathered once per pair.  Per tile, a halo pack in
// device memory is the image of the LDS region the kernel copies it to at
// launch:
//   [0, 16)        int  nh (halo rows), nset (node-set size), nblk (halo blocks), 0
//   [16, 2064)     int2 per halo row {length, first halo block}
//   [2064, 11280)  float per halo row: its 3x3 inverse diagonal block, row-major
//   [11280, 16400) ushort [u][thread]: set slot of the column of block u of each own-row lane
//   [16400, 18448) int  node of set slots 128.. (halo rows, then their other columns)
//   [18448, ..)    ushort set slot of each halo block's column (16-byte padded)
//   [.., +36 hbl)  float the 9 value components of each halo block, block after block
// followed in LDS (not in the pack) by the scratch of a pair, one float4 per
// node: the set's z, the halo rows' r and the pass values of the own and halo
// rows.
constexpr int kHaloRows = 256;       // halo rows per row tile
constexpr int kHaloSet = 640;        // node-set slots per row tile
constexpr int kHaloU = 5;            // blocks per lane of a row in pair mode (rows of at most 20 blocks)
constexpr int kHaloTileRows = 128;   // rows per tile in pair mode (4 lanes per row, 512 threads)
constexpr int kHaloOffRows = 16;
constexpr int kHaloOffInv = kHaloOffRows + 8 * kHaloRows;
constexpr int kHaloOffSlot = kHaloOffInv + 36 * kHaloRows;
constexpr int kHaloOffSet = kHaloOffSlot + 2 * kHaloU * 512;
constexpr int kHaloOffBlk = kHaloOffSet + 4 * (kHaloSet - kHaloTileRows);
__host__ __device__ constexpr long long halo_val_off(int hbl) { return kHaloOffBlk + (2LL * hbl + 15) / 16 * 16; }
__host__ __device__ constexpr long long halo_pack_bytes(int hbl) { return (halo_val_off(hbl) + 36LL * hbl + 255) / 256 * 256; }
constexpr int kHaloScratchBytes = 16 * (kHaloSet + kHaloRows + kHaloTileRows + kHaloRows);
constexpr int kHaloBaseBytes = 256;  // per-system header before the tile packs: int fail flag (nonzero: no pairs)

// Per-system descriptor for the persistent PCG kernel and the operator
// kernels.  All pointers are device pointers.  VT buffers (p, z, q) hold
// double (FP64 mode) or float (mixed mode).
// Matrix values of the block-row storage: the 9 components of 64 consecutive
// blocks form one chunk, component-major inside it ([p / 64][c][p % 64]), so a
// lane group's loads of one component are contiguous (coalesced) and the 9
// component streams of a chunk are one contiguous 4.6 KB (FP64) region.
constexpr int kValChunk = 64;
__host__ __device__ __forceinline__ size_t val_index(int p, int c) {
  return (size_t)(p >> 6) * (9 * kValChunk) + (size_t)c * kValChunk + (size_t)(p & (kValChunk - 1));
}
// entries to allocate for nnzb blocks (the last chunk padded)
__host__ __device__ __forceinline__ size_t val_count(long long nnzb) {
  return (size_t)((nnzb + kValChunk - 1) / kValChunk) * 9 * kValChunk;
}

struct SysDesc {
  int nb, n, a, k, nnzb;
  // global tile ranges
  int rt_begin, rt_end;   // row tiles (rows_per_tile block rows each)
  int et_begin, et_end;   // node tiles (PcgParams::node_tile nodes each)
  int lt_begin, lt_end;   // low-rank tiles (kLrTile W rows each)
  // chunk ranges (inclusive) of each chunking that cover this system
  int cA0, cA1;           // row tiles over the "row" workgroups (stage A, T1)
  int cF0, cF1;           // row tiles over all workgroups (Neumann stages)
  int cE0, cE1;           // node tiles over all workgroups (B, T2, init)
  int cL0, cL1;           // low-rank tiles over the low-rank workgroups
  int t_off;              // offset of this system's t vector in LDS
  int lrp_off;            // offset of this system's low-rank partials
  // matrix
  const int* rowptr;
  const int* colidx;
  const double* val64;    // component c (col-major 3x3 index) of block p at val_index(p, c)
  const float* val32;
  const double* invD64;   // [nb*9] row-major inverse diagonal blocks
  const float* invD32;
  const int* act_node;    // [nb] first W row of the node, or -1
  const int* act_dof;     // [a] FEM DOF of W row
  const double* W64;      // [a*k] row-major, V C V^T = W W^T
  const float* W32;
  // vectors
  const double* rhs;
  double* x;
  double* r;
  void* p[2];
  void* z[2];
  void* q;
  double* q64;            // FP64 D x for the true residual (== q in FP64 mode)
  const int* tstart;      // latency schedule at 4 lanes per row: row tile t covers rows [tstart[t], tstart[t+1])
                          // (at most 128, cut along D's components); null: rows [R t, R t + R)
  const int* nbr;         // [row tiles * kNbrStride] neighbour lists (local tile ids), or null
  const unsigned char* halo;  // pass pairs: kHaloBaseBytes header, then one halo pack per row tile, or null
  unsigned long long* zt;     // pass pairs: two buffers of 3 nb tagged pass values (FP32 bits | round << 32)
  double tol;
  int max_iter;
  int pad_;
};

// Row of thread group g in local row tile tl of R rows: the tile's first row
// plus g, or a row past the system (nb or more) when g is past the tile's
// end.  With a tile table (S.tstart, R = 128) the tiles follow it.
__device__ __forceinline__ int tile_row(const SysDesc& S, int tl, int R, int g) {
  if (S.tstart) {
    const int a = S.tstart[tl];
    return a + g < S.tstart[tl + 1] ? a + g : S.nb;
  }
  return tl * R + g;
}
__device__ __forceinline__ int tile_first(const SysDesc& S, int tl, int R) { return S.tstart ? S.tstart[tl] : tl * R; }
__device__ __forceinline__ int tile_end(const SysDesc& S, int tl, int R) {
  return S.tstart ? S.tstart[tl + 1] : (tl * R + R < S.nb ? tl * R + R : S.nb);
}

// Chunk tables of the streaming (strip) operator kernels: for chunk k of 64
// blocks, bit j of hbits[k] is set when block 64k + j is the first block of
// a row, and hrow[k] counts the rows that start before the chunk
// (hrow[nch] = nb).  Wave w of `nwave` walks chunks [w nch / nwave,
// (w + 1) nch / nwave); nwave 0: the launcher's choice.
struct StripDesc {
  const unsigned long long* hbits;  // [nch]
  const int* hrow;                  // [nch + 1]
  int nch, nwave;
};

// Per-system outputs of the persistent PCG kernel.
struct SysOut {
  int iterations;
  int status;             // hsc_status
  int checks;             // true-residual evaluations
  int restarts;
  int pairs;              // 1 when the launch ran its preconditioner passes in pairs
  int pad_;
  double rel;             // final relative residual (true when converged)
  double normb;
};

struct PcgParams {
  const SysDesc* sys;
  SysOut* out;
  int nsys;
  int order;
  double omegas[kMaxPolySteps];  // step j weight: z0 = w0 Dinv r; z_j = z_(j-1) + w_j Dinv (r - D z_(j-1))
  int mode_true_every;    // core.py schedule: true residual each iteration
  int replace_every;      // mixed: replace r by the FP64 true residual
  int nwg;                // grid size
  int n_lr_wg;            // workgroups [0, n_lr_wg) do the low-rank reductions
  int nA, nF, nL;         // chunk counts (nF == nwg)
  int rt_total, et_total, lt_total;
  int rows_per_tile;
  int node_tile;          // nodes per elementwise tile (<= kNodeTileMax)
  int wc_rows;            // coupled nodes per resident node tile whose W rows are cached in LDS (0 = off)
  int wc_off;             // byte offset of that cache in dynamic LDS (after the t vectors)
  const int* chunkA;      // [nwg] chunk of row-workgroup b in the A chunking (-1 for LR wgs)
  const int* chunkF;      // [nwg]
  const int* chunkL;      // [nwg] (-1 for row wgs)
  double* partA;          // [nsys * nA]
  double* partE;          // [2 * nsys * nF]  (rr, rz from node tiles)
  double* partF;          // [nsys * nF]      (rz from Neumann stages)
  double* partL;          // [sum over systems (cL1-cL0+1)*k]
  double* partR;          // pipelined loop: W^T v of each row workgroup's coupled nodes [nF * t_total]
  unsigned* barrier;      // grid barrier counter (zeroed before launch)
  int* error;             // device error flag
  long long timeout_ticks;// bounded wait (s_memrealtime ticks, 100 MHz)
  int t_total;            // LDS doubles for t vectors
  int pass_sync;          // 1: passes before the last publish per-tile flags and wait on neighbour tiles
  unsigned* pass_flags;   // [rt_total] last completed pass of each row tile (zeroed before launch)
  int sys_bytes;          // dynamic LDS bytes of the system descriptors (the low-rank node sums follow)
  int sg_bytes;           // dynamic LDS bytes of the low-rank node sums (node_tile x 3 doubles; the t vectors follow)
  int halo;               // 1: pass pairs when every system's halo packs were built
  int halo_off;           // byte offset of the halo region from the t vectors (after them and the W-row cache)
  int hbl;                // pack blocks per halo pack
  long long halo_stride;  // bytes per halo pack (halo_pack_bytes(hbl))
  long long* prof;        // optional per-workgroup stage timers [nwg * 32] (null = off):
                          // site si: [2si] ticks computing before barrier si, [2si+1] ticks waiting in it
};

// Setup-time kernels' scratch layout ------------------------------------
struct AcaParams {
  const double* E;        // qm x a row-major (input, untouched)
  double* R0;             // working copies (double buffer)
  double* R1;
  int qm, a, limit;
  double stop;            // rtol (0 = none)
  double* U;              // qm x limit, column-major
  double* V;              // a x limit, column-major
  double* hist;           // limit + 1
  int* pivots;            // limit x 2
  double* part;           // [nwg * 4]
  double* cand;           // [2 * nwg * qm] published candidate pivot columns (register-resident ACA)
  int* rank_out;
  double* scale_out;
  unsigned* barrier;
  int* error;
  long long timeout_ticks;
  int nwg;
  const int* skip;        // nonzero: the block-by-body ACA has produced the factors (exit at once)
};

// ACA by ABD body (see k_aca_blocks): scratch of one setup
struct AcaBlkParams {
  const double* E;        // qm x a row-major
  int qm, a, limit;
  double stop;            // relative-residual stop of the ACA (the body path requires 0)
  int* colbody;           // [a] body of column c, -1 none, -2 several
  double* colsq;          // [a] |E[:, c]|^2
  int* plan;              // [4 + bodies]: ok, K, ..., per-body column counts
  double* scale;          // [1] ||E||_F
  double* ublk;           // [bodies x 144] pivot columns (12 rows) of each body's steps
  double* vtmp;           // [a x 12] pivot-row entries by column and step
  double* hblk;           // [bodies x 13] residual |R_b|^2 after 0..rank steps
  int* pblk;              // [bodies x 24] pivots (global row, global column)
  int* rank_b;            // [bodies]
  // outputs in k_aca's format
  double* U;              // qm x limit column-major
  double* V;              // a x limit column-major
  double* hist;           // limit + 1
  int* pivots;            // limit x 2
  int* rank_out;
  double* scale_out;
};
constexpr int kAcaBlkMaxBodies = 128;

}  // namespace hsc
