// Kernel parameter blocks and host launchers shared by the HIP sources and
// the host orchestration.
#pragma once
#include <hip/hip_runtime.h>
#include "common.hpp"

namespace hsc {

struct SplitParams {
  const int* rows;
  const int* cols;
  const double* vals;  // col-major 3x3
  long long count;
  int qo, qm, fo, fn;
  int* row_count;      // [nb]
  int* active_flag;    // [nb]
  double* invD64;      // [nb*9] row-major
  float* invD32;
  double* A;           // [qm*qm] row-major
  int* err;            // [8]
};

struct PlaceParams {
  const int* rows;
  const int* cols;
  const double* vals;
  long long count;
  int qo, qm, fo, fn;
  const int* rowptr;
  int* cursor;
  int* key_col;
  int* key_src;        // 2*triplet + transposed
  const int* act_node;
  int a;
  double* E;           // [qm*a] row-major
};

struct AbdParams {
  const double* A;
  const double* Binv;
  double* P;        // out, qm x qm
  double* s1;       // scratch qm*qm each
  double* s2;
  double* s3;
  int qm;
  int mode;         // 0 exact, 1 neumann, 2 adaptive
  int blockdiag;    // A has no inter-body blocks
  int order;
  double omega;
  double tol;
  double* out;      // [0] residual, [1] doublings, [2] status (0 ok, 4 not SPD, 5 inverse, 6 nonfinite)
};

struct FactorParams {
  const double* P;   // qm x qm row-major
  const double* U;   // qm x ldk col-major (ld = qm)
  const double* V;   // a x ldk col-major (ld = a)
  int qm, a, k;         // k: the largest rank (buffer sizing); the rank itself is *kdev
  const int* kdev;       // rank found by ACA (device; clamped to [0, k] by the kernels)
  double* W1;        // qm x k scratch (P U)
  double* W1T;       // k x qm: (P U)^T, kept for the right-hand side of every solve
  double* C;         // k x k (out: C; then L in lower triangle of Lm)
  double* Lm;        // k x k lower factor
  double* out;       // [0] status (0 ok, 6 nonfinite / not SPD)
};

// setup_kernels.hip
hipError_t launch_split(const SplitParams& S, hipStream_t st);
hipError_t launch_scan2(const int* in0, int* out0, const int* in1, int* out1, int n, hipStream_t st);
hipError_t launch_active(const int* flag, const int* prefix, int nb, int* act_node, int* act_dof, hipStream_t st);
// *out = max over tiles of `tile` consecutive nodes of the active-node count (*out zeroed before)
hipError_t launch_max_tile_active(const int* act_node, int nb, int tile, int* out, hipStream_t st);
hipError_t launch_place(const PlaceParams& S, hipStream_t st);
hipError_t launch_sort_gather(const int* rowptr, int nb, int* key_col, int* key_src, const double* vals, int nnzb,
                              int* colidx, double* v64, float* v32, hipStream_t st);
hipError_t aca_occupancy(int* blocks);
hipError_t launch_aca(const AcaParams& P, hipStream_t st);
// ACA by ABD body when E is block diagonal by body (plan[0] = 1; else nothing)
hipError_t launch_aca_blocks(const AcaBlkParams& P, hipStream_t st);
// register-resident ACA: grid size for (qm, a) within max_wg workgroups, 0 when it does not fit
int aca_reg_grid(int qm, int a, int max_wg);
hipError_t aca_reg_occupancy(int* blocks);
hipError_t launch_aca_reg(const AcaParams& P, hipStream_t st);
hipError_t launch_body_inverse(const double* A, int qm, double* Binv, int* err, hipStream_t st);
hipError_t launch_abd_inverse(const AbdParams& S, hipStream_t st);
hipError_t launch_coupling_core(const FactorParams& S, hipStream_t st);
// W = V L for the rank *kdev (<= kmax), read on the device
hipError_t launch_make_W(const double* V, const double* Lm, int a, const int* kdev, int kmax, double* W64, float* W32,
                         hipStream_t st);
hipError_t launch_rhs(const double* W1T, const double* V, const double* b, int qo, int qm, int fo, int n, int k,
                      int a, const int* act_node, double* y, double* rhs, double* r, double* x, hipStream_t st);
hipError_t launch_backsub(const double* P, const double* U, const double* V, const int* act_dof, const double* b,
                          double* x, int qo, int qm, int fo, int a, int k, double* tmp, hipStream_t st);
hipError_t launch_full_residual(const double* yf, const double* E, const double* A, const int* act_node,
                                const int* act_dof, int a, int qm, const double* b, const double* x, int fo, int qo,
                                int n, double* part, double* out, hipStream_t st);
// pcg_kernel.hip
// res: the latency-schedule kernel (register-resident rows and node tile);
// otherwise the throughput-schedule kernel (streaming SpMV with several
// blocks per lane in flight)
// pipe: the pipelined loop (one grid barrier per iteration; G = 4, latency schedule only)
hipError_t launch_pcg(int G, bool mixed, bool res, bool pipe, const PcgParams& P, size_t shmem, hipStream_t st);
hipError_t pcg_occupancy(int G, bool mixed, bool res, bool pipe, size_t shmem, int* blocks);
hipError_t launch_pcg_plain(int G, bool mixed, bool res, bool pipe, const PcgParams& P, size_t shmem, hipStream_t st);
const void* pcg_kernel_fn(int G, bool mixed, bool res, bool pipe);
hipError_t launch_zero(void* p, size_t bytes, hipStream_t st);  // bytes % 4 == 0; a kernel, safe in graphs
// per-tet log-Neo-Hookean Hessian (SPD-projected when project != 0) pulled
// back to 10 node-block triplets per tet (fem_hessian.hip)
struct FemHessianParams {
  const int* tets;        // [ntet][4] node indices
  const double* x;        // [nodes][3] positions
  const double* Dm_inv;   // [ntet][9] column-major inverse rest edge matrices
  const double* mu;       // [ntet] Lame parameters
  const double* lambda;
  const double* vol;      // [ntet] rest volumes
  double dt2;             // scale dt^2 (with the volume)
  long long ntet;
  int project;
  int* orow;              // [10 ntet] block rows (the smaller node index)
  int* ocol;
  double* oval;           // [10 ntet][9] column-major blocks
};
hipError_t launch_fem_hessian(const FemHessianParams& P, hipStream_t st);
struct AbdFemCouplingParams {
  const int* abd_vertex;    // [count] ABD vertex per contact block
  const int* fem_vertex;    // [count] FEM vertex per contact block
  const double* hess;       // [count][9] column-major 3x3 (ABD rows, FEM columns)
  const int* v2b;           // [n_abd_vertices] body of each ABD vertex
  const double* xbar;       // [n_abd_vertices][3] rest positions
  const int* body_fixed;    // [n_bodies]
  const int* vertex_fixed;  // [n_fem_vertices]
  long long count;
  int n_abd_vertices, n_fem_vertices, n_bodies;
  int abd_block_offset, fem_block_offset;
  int* err;                 // set to 1 on an index out of range
  int* orow;                // [4 count]
  int* ocol;
  double* oval;             // [4 count][9]
};
// launch only (no synchronisation), assemble.hip
hipError_t launch_abd_fem_coupling(const AbdFemCouplingParams& P, hipStream_t st);
// general 3x3-block triplets -> upper-triangle BCOO of unique blocks, sorted by
// (row, col), duplicates summed in input order (assemble.hip); synchronises st
// (workspace: library-held, one per device, grown on demand; counted by note_alloc)
hipError_t assemble_bcoo(const int* rows, const int* cols, const double* vals, long long count, int fold,
                         int* orow, int* ocol, double* ovals, long long* out_count, hipStream_t st);
// workspace layout of one assembly of n triplets (hipcub's size queries only:
// no allocation) and the assembly itself in a caller-provided workspace of
// at least total bytes (nothing in it needs to start at zero)
struct AsmLayout {
  size_t kb = 0, ib = 0, tmp = 0, total = 0;
  size_t sort_bytes = 0, scan_bytes = 0;
};
hipError_t asm_layout(int n, hipStream_t st, AsmLayout* L);
hipError_t assemble_in(const int* rows, const int* cols, const double* vals, long long count, int fold, int* orow,
                       int* ocol, double* ovals, long long* out_count, char* arena, const AsmLayout& L,
                       hipStream_t st);
// one launch of a row-mapped operator kernel (G lanes per block row): 0 FP64 D x,
// 1 FP64 pass, 2 FP32 pass (y = x + w Dinv (r - D x), r = S.r)
hipError_t launch_row_op(int kernel, int G, const SysDesc& S, const void* x, void* y, double w, hipStream_t st);
// up to kZeroListMax buffers zeroed by one kernel launch (each bytes % 4 == 0)
constexpr int kZeroListMax = 12;
struct ZeroList {
  void* p[kZeroListMax];
  long long start[kZeroListMax + 1];
  int n;
};
hipError_t launch_zero_multi(const void* const* ptrs, const size_t* bytes, int n, hipStream_t st);
// up to kGatherMax small copies device -> device by one launch (n words of
// 4 bytes each; a null src writes zeros)
constexpr int kGatherMax = 12;
struct GatherList {
  const unsigned* src[kGatherMax];
  unsigned* dst[kGatherMax];
  int n[kGatherMax];
  int count;
};
hipError_t launch_gather_words(const GatherList& L, hipStream_t st);
// streaming (strip) operator kernels: chunk tables, and one launch of
// 0 FP64 D x, 1 FP64 pass, 2 FP32 pass (y = x + w Dinv (r - D x))
hipError_t launch_chunk_heads(const int* rowptr, int nb, int nnzb, unsigned long long* hbits, int* hrow,
                              hipStream_t st);
hipError_t launch_strip_op(int kernel, const SysDesc& S, const StripDesc& T, const void* x, void* y, double w,
                           hipStream_t st);
// the preconditioner's passes, FP64: streaming pass kernels (T) or row-mapped
// ones with G lanes per row (T null)
hipError_t launch_op_precond64(const StripDesc* T, int G, const SysDesc& S, int order, const double* omegas,
                               const double* r, double* z, double* tmp, hipStream_t st);
hipError_t launch_hash_fill(double* x, int n, unsigned seed, hipStream_t st);
// body-aligned row tiles: tstart (row_tiles_capacity() ints) gets the tile
// starts and nb, out[0] the tile count (0: none, the fixed tiles apply),
// out[1] the most coupled nodes in one tile (reach: nb ints of scratch)
hipError_t launch_row_tiles(const int* rowptr, const int* colidx, const int* act_node, int nb, int* reach,
                            int* tstart, int* out, hipStream_t st);
int row_tiles_capacity();
// row-tile neighbour lists for R rows per tile, or for the ntiles tiles of a
// tile table (bits: ceil(rtl/32)*rtl words, list: rtl*kNbrStride ints,
// rtl = ceil(nb/R) or ntiles, at most kNbrMaxTiles)
hipError_t launch_tile_neighbors(const int* rowptr, const int* colidx, int nb, int R, const int* tstart, int ntiles,
                                 unsigned* bits, int* list, hipStream_t st);
// pass pairs of a system (128-row tiles, or the ntiles tiles of a tile
// table): the halo packs (kHaloBaseBytes header with the fail word, then
// halo_pack_bytes(hbl) per tile) into halo
hipError_t launch_tile_pairs(const int* rowptr, const int* colidx, const float* val32, const float* invD32, int nb,
                             const int* tstart, int ntiles, int hbl, unsigned char* halo, hipStream_t st);
// static LDS bytes of the persistent PCG kernel variant
hipError_t pcg_static_lds(int G, bool mixed, bool res, bool pipe, size_t* bytes);
// power iteration: z = Dinv D x / |x| with |x|^2 the sum of part_in[0:np_in];
// part gets power_step_tiles(G, nb) partials of |z|^2
int power_step_tiles(int G, int nb);
hipError_t launch_power_step(int G, const SysDesc& S, const double* x, const double* part_in, int np_in, double* z,
                             double* part, hipStream_t st);
// fixed-order sum of np partials into *out
hipError_t launch_sum_partials(const double* part, int np, double* out, hipStream_t st);
hipError_t launch_op_lowrank64(const SysDesc& S, const double* x, double* y, double* t, hipStream_t st);

// partials of a.b only (*np_out <= 256 of them, into part)
hipError_t launch_dot_partial(const double* a, const double* b, int n, double* part, int* np_out, hipStream_t st);
hipError_t launch_node_graph(const int* rowptr, const int* colidx, const double* val64, int nb, int* nbr_ptr,
                             int* nbr, float* Dfeat, float* sdiag, hipStream_t st);

}  // namespace hsc
