// fcp_internal.h -- context, device buffers and kernel entry points of libfcp.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/fcp.h"

// ---------------------------------------------------------------------------
// Reference literals (parameters.f90:6, linear_solvers.f90:276,447): default
// REAL literals widened to REAL(8).
// ---------------------------------------------------------------------------
#define FCP_SMALL ((double)1e-20f)
#define FCP_S0INIT ((double)1e20f)

// Sweep outputs are pre-filled with a signalling-NaN pattern that arithmetic
// never produces; a consumer spins until the producer overwrote it.
#define FCP_SENT64 0x7FF4DEAD5EED0001ull
#define FCP_SENT32 0x7FA5EED1u

// Bound of one sweep dependency wait in ticks of the 100 MHz wall clock
// (2 s), after which the wave records FCP_ERR_SWEEP and moves on.
#define FCP_WAIT_TICKS 200000000ull

template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  hipError_t alloc(size_t count) {
    if (count <= n && p) return hipSuccess;
    release();
    n = count;
    if (count == 0) return hipSuccess;
    return hipMalloc(&p, count * sizeof(T));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DBuf() { release(); }
};

// Level schedule of one triangular sweep: rows in processing order grouped
// into tasks of <= 64 rows that share a level (one wavefront per task).
struct Schedule {
  int nlev = 0, ntasks = 0;
  DBuf<int> order;       // [n] rows in level order
  DBuf<int> task_ptr;    // [ntasks+1] offsets into order
  DBuf<int> lev;         // [n] level of each row
  std::vector<int> lev_ptr;  // [nlev+1] host offsets of every level in order
};

// Row-block schedule of one triangular sweep: contiguous blocks of R rows,
// each block's rows ordered by level ("steps"); one workgroup per block keeps
// the block's results in LDS and waits only for rows of earlier blocks.
struct BlockSchedule {
  int R = 0, nblocks = 0, nsteps = 0, max_steps = 0;
  bool tiled = false;  // blocks are geometric tiles (else contiguous rows)
  DBuf<int> rows;      // [n] block-major, level-minor, ascending row inside a level
  DBuf<int> step_ptr;  // [nsteps+1] offsets into rows
  DBuf<int> blk_step;  // [nblocks+1] first step of every block
  DBuf<int> blk_pos;   // [nblocks+1] first position (index in rows) of every block
  // Packed step-major operands for the pipelined sweep, position = index in
  // `rows`: dependency count, dependency (>= 0: position inside the block =
  // its LDS slot, < 0: -(row+1) of an earlier block), CSR position of its
  // coefficient.
  int maxd = 0;
  DBuf<int> prc;                // [2*n]  (row, count) pairs
  DBuf<int> pdep;               // [PMAXD*n] dependencies, PMAXD per position
  DBuf<int> psrc;               // [PMAXD*n] CSR positions of the coefficients
  DBuf<double> pa64, pdx64;     // [PMAXD*n] coefficients; [2*n] (pivot, spare)
  DBuf<float> pa32, pdx32;
  uint64_t pver64 = ~0ull, pver32 = ~0ull;
};
#define FCP_PMAXD 4  // dependencies per row handled by the pipelined sweep
#define FCP_COARSE_MAXN 2048  // most tiles that get the inner PCG's coarse correction

// Row-block schedules of a batched (several right-hand sides) sweep, whose
// block size depends on the bytes per row kept in LDS.
struct MultiSched {
  BlockSchedule lo, up;
  DBuf<char> pin;  // packed sweep inputs, one row of values per position
};

// Device-resident scalars of one Krylov solve (one column).
struct SolveScalars {
  double sk, s0, pkapk, alf, bet, res0, resl, factor, resor, rsm;
  double l2, l2b, om, gam, beto, ukreso, vv, vr;
  double tol;  // per-column relative tolerance of a batched inner solve
  int iter, done, status, conv;  // conv: the stopping test was met
  int imax;                      // iteration limit of a tile-order inner solve
};

// Which residual ratio a polling loop watches: the reference's L1 ratio rsm
// against tol_rel, the L2 ratio l2/l2b against tol_rel, or the L2 ratio
// against each column's own inner tolerance S.tol.
enum { BATCH_L1 = 0, BATCH_L2 = 1, BATCH_INNER = 2 };

// Length of the next batch of iterations launched between two host polls of
// the device scalars.  At each poll, every unfinished column's residual ratio
// and iteration count are compared with the previous poll; the geometric rate
// per iteration since then predicts how many more iterations the column needs
// to meet its stopping test.  The batch covers the slowest column plus one
// iteration (an iteration's stopping test runs in the first kernel of the
// next), within [1, cap]; without a usable rate (first poll, a residual that
// did not decrease) it is cap.  Iterations launched after the device set
// `done` do no work, so a short estimate costs one more poll and a long one
// only launches that exit at once.
struct BatchSizer {
  static constexpr int MAXC = 8;
  int cap, mode;
  int last_need = 0;  // the last next()'s estimate before the cap (cap while no rate is known)
  double tol;
  double rho[MAXC];
  int it[MAXC];
  BatchSizer(int cap_, int mode_, double tol_) : cap(cap_ > 0 ? cap_ : 8), mode(mode_), tol(tol_) {
    for (int c = 0; c < MAXC; ++c) {
      rho[c] = 0.0;
      it[c] = -1;
    }
  }
  // Batch length after a poll of ncol columns' host copies h.
  int next(const SolveScalars *h, int ncol) {
    int out = 1;
    for (int c = 0; c < ncol && c < MAXC; ++c) {
      const SolveScalars &S = h[c];
      if (S.done) continue;
      const double ratio = mode == BATCH_L1 ? S.rsm : (S.l2b > 0.0 ? S.l2 / S.l2b : 0.0);
      const double target = mode == BATCH_INNER ? S.tol : tol;
      int est = cap;
      if (it[c] >= 0 && S.iter > it[c] && ratio > 0.0 && ratio < rho[c] && target > 0.0) {
        const double q = log(ratio / rho[c]) / (double)(S.iter - it[c]);
        const double need = target < ratio ? log(target / ratio) / q : 0.0;
        est = need < 1e6 ? (int)ceil(need) + 1 : 1000000;
      }
      rho[c] = ratio;
      it[c] = S.iter;
      if (est > out) out = est;
    }
    last_need = out;
    return out > cap ? cap : out;
  }
};

struct ProfEntry {
  int launches = 0;
  double total_ms = 0.0;
};

struct fcp_ctx {
  int device = 0;
  int ncu = 0;
  hipStream_t stream = nullptr;
  std::string err;

  // -------- mesh (0-based, resident) --------
  bool have_mesh = false;
  int n = 0, F = 0, Fi = 0, Fb = 0, nNodes = 0, nPatches = 0;
  std::vector<int> h_nfaces, h_start, h_bctype;
  DBuf<double> points;  // [3*nNodes]
  DBuf<int> nnodes;     // [F]
  DBuf<int> node;       // [NOMAX*F]
  DBuf<int> owner;      // [F]
  DBuf<int> neighbour;  // [Fi]
  DBuf<int> bpatch;     // [Fb] patch of every boundary face
  DBuf<int> bctype;     // [nPatches]
  DBuf<int> cf_ptr;     // [n+1] cell -> faces (ascending face id)
  DBuf<int> cf_face;    // [2*Fi+Fb]
  // periodic patches (geometry.f90:251-258, 1848-1997): the twin start face
  // of every periodic patch, the owners as uploaded (face mapping rewrites
  // owner and the twin faces' geometry), and per periodic pair k (patch
  // order): its face, its twin face, the face's index in its patch; per
  // boundary face the pair it belongs to (kof) or whose average it takes
  // last (twb); per cell the pairs whose twin face it owns (tw_ptr / tw_k)
  std::vector<int> h_twin;
  DBuf<int> owner0;
  int nper = 0;
  DBuf<int> per_kof, per_twb, per_fp, per_ft, per_ii, per_twptr, per_twk;

  // -------- geometry --------
  bool have_geom = false;
  DBuf<double> xc, yc, zc, vol;
  DBuf<double> arx, ary, arz, xf, yf, zf;
  DBuf<double> facint, Df, dnb, srdb;

  // -------- CSR --------
  bool have_csr = false;
  int nnz = 0;
  DBuf<int> ia, ja, diag, slot_pn, slot_np;
  DBuf<int> row_of;     // [nnz] row of every entry
  Schedule lower, upper;
  BlockSchedule blower, bupper;      // FP64-sized blocks
  BlockSchedule blower4, bupper4;    // FP32-sized blocks (twice the rows)
  BlockSchedule tlower, tupper;      // geometric-tile blocks (when acyclic)
  bool have_tiles = false;
  // tile-local preconditioner of the FP32 inner solves: geometric tiles (any
  // numbering) or contiguous blocks, dependencies inside the block only
  BlockSchedule jlower, jupper;
  bool have_bj = false;
  bool bj_wave = false;              // tiles small enough for one wave each (steps <= 64 rows)
  std::vector<int> bj_tile;          // [n] block of every row (host)
  DBuf<double> dbj;                  // [n] tile-local DIC / DILU pivots (inverted)
  uint64_t bj_version = ~0ull, bj_epoch = 0;
  int bj_kind = -1;
  int inner_precond = 1;             // 0 global DIC/DILU, 1 tile-local in tile order,
                                     // 2 tile-local with row-gather sweeps
  // FP32 inner PCG in the tile-local preconditioner's processing order
  bool tin_have = false;             // pattern of the current tile-local schedule
  uint64_t tin_epoch = ~0ull;        // pivot epoch of the packed values
  DBuf<int> tin_iperm;               // [n] position of every row
  DBuf<int> tin_dl, tin_du;          // [2n] 4 x 16-bit LDS slots of both sweeps
  DBuf<float> tin_al, tin_au, tin_dv;  // [4n] coefficients, [n] pivots
  DBuf<int> tin_base, tin_col;       // SELL-64 copy of P A P^T
  int tin_width = 0;                 // its uniform slice width (0: widths differ)
  DBuf<float> tin_val;
  // packed sweep layout: three 10-bit dependency distances per position and
  // sweep in the bits of the coefficients' fourth component (every position
  // with at most 3 in-tile dependencies per sweep, distances below 1024)
  bool tin_pack = false;
  DBuf<unsigned> tin_offl, tin_offu;
  int tune_tile_pack = 1;            // FCP_TUNE_TILE_PACK: use the packed layout where it applies
  int tune_coarse_fuse = 1;          // FCP_TUNE_COARSE_FUSE: coarse correction inside tk_pre
  DBuf<unsigned> tin_bar;            // grid barrier of the fused tk_pre (two 128-byte lines)
  // coarse correction of the inner PCG (one unknown per tile): tile of every
  // position, tile adjacency, the coarse operator (then its factor, then its
  // inverse; ncp x ncp row-major) and the inverse factor
  int inner_coarse = 1;              // add the coarse correction when it applies
  int tc_nc = 0, tc_ncp = 0, tc_maxdeg = 0;  // tiles (0: none), padded to 64, widest adjacency
  bool tc_have = false;              // adjacency of the current tile schedule examined
  uint64_t tc_epoch = ~0ull;         // matrix version of the factor
  DBuf<int> tin_usrc;                // CSR entry of every in-tile upper sweep slot (-1: none)
  DBuf<int> tin_rs;                  // CSR start and length of every position's row
  DBuf<int> tc_tpos, tc_cptr, tc_cadj, tc_flag;  // tc_flag: the factor failed
  DBuf<unsigned char> tc_eslot;      // adjacency slot of every CSR entry
  hipStream_t tc_stream = nullptr;   // side stream of the coarse factorisation
  hipEvent_t tc_ready = nullptr, tc_done = nullptr;
  bool tc_pending = false;           // a factorisation on tc_stream not yet joined
  DBuf<double> tc_L, tc_W, tc_r0;
  DBuf<float> tin_vec[7];            // PCG: x, r, z, q, p (two); BiCGStab: x, r, z, p, u, v, reso
  // FP64 copies of the tile-order operands and PCG vectors (FCP_PREC_FP64_TILE)
  uint64_t tin_epoch64 = ~0ull;      // pivot epoch of the FP64 values
  DBuf<double> tin_al64, tin_au64, tin_dv64, tin_val64;
  DBuf<double> tin_vec64[6];
  bool bj_active = false;            // FP32 sweeps use the tile-local preconditioner
  DBuf<int> tslot;      // [nnz] position of the transposed entry (DILU)
  bool sym_pattern = true;  // every entry has its transposed entry
  // SELL-64 copy of the operator for SpMV: rows in slices of 64, each slice
  // stored column-major (entry k of the slice's rows contiguous) and padded
  // to the slice's longest row; per-row entry order = CSR order.
  int nslices = 0;
  DBuf<int> sell_base;  // [nslices+1] offset of each slice (multiple of 64)
  DBuf<int> sell_col;   // [sell_base[nslices]]
  int sell_width = 0;   // uniform slice width (0: widths differ, read sell_base)
  DBuf<double> sell_a64;
  DBuf<float> sell_a32;
  uint64_t sell64_version = ~0ull, sell32_version = ~0ull;

  // -------- operator values --------
  DBuf<double> a;       // [nnz] resident matrix (fp64)
  DBuf<float> a32;      // [nnz] fp32 shadow
  uint64_t a_version = 0, a32_version = ~0ull, dic_version = ~0ull, d32_version = ~0ull;
  int dic_kind = -1;    // 0 DIC, 1 DILU
  uint64_t d_epoch = 0; // bumped at every preconditioner setup

  // -------- solver workspace --------
  DBuf<double> d, res, pk, zk, yk, uk, vk, reso, x, rhs, qk;
  DBuf<float> d32, r32, p32, z32, y32, q32, e32, u32, v32, reso32;
  DBuf<double> partials;     // per-block partial sums
  DBuf<int> counters;        // [16] block-completion counters / tickets
  DBuf<SolveScalars> scal;   // device scalars
  SolveScalars *h_scal = nullptr;  // pinned mirror
  int *h_err = nullptr;            // pinned mirror of errflag
  // the tile-order PCG's pipelined polls: two pinned slots of 4 scalars and
  // an error flag each, and the event that marks each slot's copy complete
  SolveScalars *h_poll = nullptr;
  int *h_perr = nullptr;
  hipEvent_t poll_ev[2] = {nullptr, nullptr};
  DBuf<int> errflag;         // [1]

  // -------- restarted GMRES with ILU(0) --------
  DBuf<double> ilu;       // [nnz] ILU(0) factors of `a` (ilu_cr layout)
  uint64_t ilu_version = ~0ull;
  DBuf<double> gm_v;      // [(mr+1)*n] Krylov basis, column-major
  DBuf<char> gm_state;    // device GMRES scalars and Hessenberg matrix
  int gm_restart = 4;     // restart length of solver "pmgmres" (csrsolve :85)
  // the LIS-replacement solver (LIS_linear_solvers.f90:38-200): the options
  // string and what it selects (solver code of fcp_csrsolve, GMRES restart,
  // tolerance, iteration limit, zero initial guess, convergence condition)
  struct LisSolver {
    bool have = false;
    std::string options;
    int code = 0, restart = 40, maxiter = 1000, conv_cond = 0;
    double tol = 1e-12;
    bool initx_zeros = true;
  } lis;

  // -------- least-squares gradient matrices --------
  DBuf<double> lsq_dmat;  // Dmat(9,n) or D(3,6,n)
  DBuf<double> fvx_old;   // [3*(n+Fb)] old gradient of the fvx Gauss sweeps
  DBuf<double> cd_gamf;   // [F] face diffusivity of fcp_convdiff_cell
  DBuf<double> pc_fac;    // [1] adjustMassFlow's outlet scaling factor
  DBuf<double> pc_s;      // [3n] su, sv, sw of fcp_pressure_correct when not asked for
  DBuf<int> mom_wall;     // [Fb] wall index of every boundary face (-1: not a wall)
  // cell blocks of the per-cell assembly kernels (256 cells each): the inner
  // faces a block's cells own ([asm_fo[b], asm_fo[b+1]), inner faces sorted
  // by owner) and the faces reaching it from other blocks (asm_inc), staged in
  // LDS; per cell-face incidence the face's LDS slot (bit 30: the cell is the
  // face's neighbour).  asm_state 0 not built for this CSR, 1 ready, -1 not
  // applicable (the per-cell kernels read faces from HBM instead).
  int asm_state = 0;
  bool asm_rows = false;             // block rows fit the LDS row buffer
  int asm_nb = 0;
  DBuf<int> asm_fo, asm_incp, asm_inc, asm_slot;
  int tune_asm_blk = 1;              // FCP_TUNE_ASM_BLOCKS
  DBuf<double> mom_grad, mom_s3, mom_dp, mom_w3[5];  // momentum assembly scratch
  int lsq_method = 0, lsq_compat = 1;

  // -------- scalar transport step (calculate_temperature) --------
  DBuf<unsigned char> tr_kind;  // [Fb] boundary treatment per face
  DBuf<int> tr_srd;             // [Fb] face whose |S|/dn a wall term reads
  std::vector<double> tr_wf;    // wall fractions the two above were built for
  DBuf<double> tr_grad, tr_dcoef, tr_su, tr_sp, tr_mm;

  // -------- batched solves (interleaved vectors, NRP values per row) --------
  std::map<long long, std::unique_ptr<MultiSched>> msched;  // keyed by rows per block (lo, up)
  DBuf<char> tpin;  // batched sweep inputs on the tile schedules
  DBuf<char> jpin;  // batched sweep inputs on the tile-local schedules
  DBuf<double> mv64[9];
  DBuf<float> mv32[9];
  // per-column diagonals of a batched solve (velocity.f90:611-632: the
  // momentum components share the off-diagonals): the diagonals and the
  // pivots interleaved by row, the matrix with one column's diagonal (pivot
  // set-up), and the pivots in each sweep schedule's processing order
  bool mdc_active = false;
  DBuf<double> mdcol, mpiv, mdiag_a;
  DBuf<double> mpdc[2];
  const void *mpdc_of[2] = {nullptr, nullptr};

  // -------- staging --------
  DBuf<char> stage[4];

  // -------- launch shapes --------
  int tune_sweep_waves = 0;
  int tune_sweep_sleep = 1;
  int tune_sweep_kind = 0;     // 0 row-block (LDS) sweeps, 1 level-task sweeps
  int tune_sweep_rows = 0;     // rows per block of the row-block sweeps (0 = auto)
  int tune_sweep_tile = 0;     // rows per geometric tile (0 = auto, 1 = no tiles)
  int tune_inner_tile = 0;     // rows per tile of the tile-local preconditioner (0 = auto)
  int tune_graph = 0;          // inner iterations replayed from a captured HIP graph
  int tune_spmv = -1;          // standalone SELL SpMV variant (FCP_TUNE_SPMV; -1 automatic)
  // captured pair of tile-order inner iterations and the launch arguments it holds
  hipGraphExec_t tin_graph = nullptr;
  std::vector<const void *> tin_graph_key;
  size_t lds_cap = 65536;      // LDS bytes available to one workgroup
  int block_rows_cap[2] = {65536, 65536};  // rows per contiguous block (lower, upper sweep)

  // -------- profiling --------
  bool prof = false;
  std::map<std::string, ProfEntry> prof_tab;
  struct PendingEv {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<PendingEv> prof_pending;
  std::vector<hipEvent_t> ev_pool;
};

// ---------------------------------------------------------------------------
// helpers (fcp_api.hip)
// ---------------------------------------------------------------------------
int fcp_fail(fcp_ctx *ctx, int code, const char *fmt, ...);
#define FCP_HIP(ctx, x)                                                          \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess)                                                        \
      return fcp_fail(ctx, FCP_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #x, \
                      hipGetErrorString(e_));                                    \
  } while (0)
#define FCP_TRY(x)            \
  do {                        \
    int rc_ = (x);            \
    if (rc_ < 0) return rc_;  \
  } while (0)

bool fcp_is_device_ptr(const void *p);

// Profiling brackets around a launch on ctx->stream.
void fcp_prof_begin(fcp_ctx *ctx, const char *name, hipEvent_t *ev);
void fcp_prof_end(fcp_ctx *ctx, const char *name, hipEvent_t ev);
struct ProfScope {
  fcp_ctx *ctx;
  const char *name;
  hipEvent_t ev = nullptr;
  ProfScope(fcp_ctx *c, const char *nm) : ctx(c), name(nm) { fcp_prof_begin(ctx, name, &ev); }
  ~ProfScope() { fcp_prof_end(ctx, name, ev); }
};

inline int grid_for(long long n, int block) { return (int)((n + block - 1) / block); }

namespace fcpk {
const char *noalloc_message();  // fcp_noalloc.hip
}

// ---------------------------------------------------------------------------
// module entry points (host-side launchers)
// ---------------------------------------------------------------------------
// fcp_mesh.hip
int fcp_build_cell_faces(fcp_ctx *ctx);
int fcp_compute_geometry(fcp_ctx *ctx);
// Device view of the periodic pairs (np = 0 and null arrays without them).
struct PerP {
  int np;
  const int *kof, *twb, *fp, *ft, *ii, *tw_ptr, *tw_k;
};
PerP fcp_periodic_view(const fcp_ctx *ctx);
int fcp_build_csr(fcp_ctx *ctx);
int fcp_build_schedules(fcp_ctx *ctx);
int fcp_build_block_schedules(fcp_ctx *ctx);
// Rows per block for elements of `elem` bytes: the tuned value (FP64 rows;
// doubled for FP32), or the largest multiple of 64 (<= 65536) whose block
// fits the workgroup LDS limit with 4 KB left for the step table.
// The automatic value is further capped per sweep direction (upper = 1) by
// the block size whose level steps pipeline best (chosen at schedule build).
inline int fcp_block_rows(const fcp_ctx *ctx, int elem = 8, int upper = 0) {
  if (ctx->tune_sweep_rows > 0) {
    long long t = (long long)ctx->tune_sweep_rows * 8 / elem;
    t = (t / 64) * 64;
    return (int)(t < 64 ? 64 : t);
  }
  long long r = ((long long)ctx->lds_cap - 4096) / elem;
  r = (r / 64) * 64;
  if (r > 65536) r = 65536;
  if (r > ctx->block_rows_cap[upper ? 1 : 0]) r = ctx->block_rows_cap[upper ? 1 : 0];
  if (r < 64) r = 64;
  return (int)r;
}
// fcp_mesh.hip: row-block schedule of one triangle (upper = 1) with R rows
// per block, packed static operands included.
// tile_rank (host, may be NULL): block rank of every row; blocks are then
// the tiles in rank order (R = the largest tile) instead of row / R.
// local_only: dependencies in other blocks are dropped (tile-local
// preconditioner) and no row is external.
int fcp_build_block_schedule(fcp_ctx *ctx, const Schedule &s, int R, int upper, BlockSchedule &b,
                             const int *tile_rank = nullptr, int ntiles = 0,
                             bool local_only = false, int step_max = 256);
int fcp_build_bj_schedules(fcp_ctx *ctx);
int fcp_build_tile_schedules(fcp_ctx *ctx);
// fcp_assemble.hip
int fcp_launch_laplacian(fcp_ctx *ctx, const double *mu, const double *mu_f,
                         const unsigned char *bc_kind, const double *phi, double *su);
int fcp_launch_nonorth(fcp_ctx *ctx, const double *gamma_f, const double *grad, double *su);
int fcp_launch_grad_gauss(fcp_ctx *ctx, const double *u, double *dudxi);
// fvxGradient.f90 Gauss gradient: dfo == nullptr runs grad_gauss's two
// sweeps (the first into tmp[3*(n+Fb)]), else grad_gauss_corrected's one
// sweep from dfo[3*n] (which must not alias dudxi).
int fcp_launch_grad_fvx(fcp_ctx *ctx, const double *u, const double *dfo, double *tmp,
                        double *dudxi);
int fcp_launch_update_boundary(fcp_ctx *ctx, double *phi);
// facefluxscalar's face diffusivity gf[F] from the cell array dc[n].
int fcp_launch_face_gamma_cells(fcp_ctx *ctx, const double *dc, double *gf);
// Pressure correction of SIMPLE (calcp_simple.f90:69-419) on the resident
// matrix: the assembly (a, su; flmass and, with adjustMassFlow, the outlet
// values of u, v, w updated) and the correction from the solved pp.
int fcp_launch_pcorr_assemble(fcp_ctx *ctx, double *u, double *v, double *w, const double *p,
                              const double *dpdxi, const double *den, const double *apu,
                              const double *apv, const double *apw, double *flmass,
                              int const_mflux, double flomas, double *su);
int fcp_launch_pcorr_correct(fcp_ctx *ctx, double *pp, double *flmass, double *u, double *v,
                             double *w, double *p, double *dpdxi, const double *apu,
                             const double *apv, const double *apw, double urfp, int pref,
                             double *su, double *sv, double *sw);
// calcuvw's assembly (velocity.f90:50-750) before its solves: u, v, w and p
// boundary values updated, dpdxi [3 nt], grad9 [9 nt] (u, v, w gradients),
// s3 [3n] scratch; rhs3, diag3, ap3, sp3, src3 [3n] per component.
int fcp_launch_momentum(fcp_ctx *ctx, double *u, double *v, double *w, const double *uo,
                        const double *vo, const double *wo, const double *uoo,
                        const double *voo, const double *woo, double *p, const double *den,
                        const double *vis, const double *visw, const int *wall_of,
                        const double *flmass, double viscos, double dt, int ts, int scheme,
                        double gds, const double *urf, double *dpdxi, double *grad9,
                        double *s3, double *rhs3, double *diag3, double *ap3, double *sp3,
                        double *src3);
int fcp_launch_convdiff(fcp_ctx *ctx, const double *flmass, const double *dcoef, double *phi,
                        const double *grad, double gam, int scheme, const unsigned char *bc_kind,
                        double urf, double *su, double *sp, const int *srd_map = nullptr,
                        int cn = 0, const double *den = nullptr, double dt = 1.0,
                        const double *phi_o = nullptr);
int fcp_launch_transport_dcoef(fcp_ctx *ctx, const double *vis, const unsigned char *bc_kind,
                               double viscos, double pranl, double sigt, double *dcoef);
// mm[2*1024] partials; mm[0], mm[1] receive min and max of x[0..n)
int fcp_launch_minmax_clip(fcp_ctx *ctx, double *x, int clip, double *mm);
int fcp_launch_ddt(fcp_ctx *ctx, int scheme, const double *den, double dt, const double *fo,
                   const double *foo, double *su, double *sp);
// fcp_adjoint.hip
int fcp_launch_transpose(fcp_ctx *ctx);
int fcp_launch_convdiff_bwd(fcp_ctx *ctx, const double *flmass, const double *dcoef,
                            const double *phi_old, const double *grad_old, double gam, int scheme,
                            const unsigned char *bc_kind, const double *lam,
                            const double *phi_new, double *dcoef_bar, double *phi_old_bar,
                            double *grad_old_bar);
int fcp_launch_laplacian_bwd(fcp_ctx *ctx, const double *mu_f, const unsigned char *bc_kind,
                             const double *phi_bc, const double *gamma, const double *grad,
                             const double *lam, const double *phi, double *mu_f_bar,
                             double *phi_bc_bar, double *gamma_bar, double *grad_bar);
int fcp_launch_grad_gauss_bwd(fcp_ctx *ctx, const double *grad_bar, double *phi_bar);
// fcp_lsq.hip
int fcp_launch_lsq_matrix(fcp_ctx *ctx, int method, int compat);
int fcp_launch_grad_lsq(fcp_ctx *ctx, const double *phi, double *dphi);
// fcp_krylov.hip
int fcp_build_sell(fcp_ctx *ctx);
int sell_uniform(fcp_ctx *ctx, int *len, int nslices, int *width);
int fcp_launch_spmv64(fcp_ctx *ctx, const double *a, const double *x, double *y);
int fcp_solve_device(fcp_ctx *ctx, int solver, double *x, const double *b, int itr_max,
                     double tol_abs, double tol_rel, int verbose, fcp_solve_opts *opts);
int fcp_ensure_workspace(fcp_ctx *ctx);
int fcp_precond_device(fcp_ctx *ctx, int kind, const double *r, double *z, double *d_out);
int fcp_bj_device(fcp_ctx *ctx, int kind, const double *r, double *z, double *d_out);
int fcp_bj32_device(fcp_ctx *ctx, int kind, const double *r, double *z, double *d_out,
                    bool fp64 = false);
// fcp_inner.hip: a pending coarse factorisation of the inner PCG (side
// stream) joined before the matrix changes / started after a Laplacian
// assembly
int fcp_coarse_join(fcp_ctx *ctx);
int fcp_coarse_wait(fcp_ctx *ctx);
int fcp_coarse_begin(fcp_ctx *ctx);
// fcp_gmres.hip: pmgmres_ilu (restart mr) on the resident matrix; ILU(0)
// factors; y = A^T x
int fcp_gmres_device(fcp_ctx *ctx, double *x, const double *b, int itr_max, int mr,
                     double tol_abs, double tol_rel, int verbose, fcp_solve_opts *opts);
int fcp_ilu_factor_device(fcp_ctx *ctx);
int fcp_ilu_apply_device(fcp_ctx *ctx, const double *r, double *z);
int fcp_atx_device(fcp_ctx *ctx, const double *x, double *y);
// fcp_multi.hip: batched solve of nrhs columns (column-major, device memory)
int fcp_solve_multi_device(fcp_ctx *ctx, int solver, int nrhs, double *fi, int ldfi,
                           const double *rhs, double *res0_out, int itr_max, double tol_abs,
                           double tol_rel, int verbose, fcp_solve_opts *opts,
                           const double *diagv = nullptr);

// rocPRIM-based primitives (fcp_prim.hip)
int fcp_exclusive_scan_int(fcp_ctx *ctx, const int *in, int *out, int n);
int fcp_sort_keys_u64(fcp_ctx *ctx, const unsigned long long *keys_in,
                      unsigned long long *keys_out, int n, int end_bit);
// fcp_rcm.hip: reverse Cuthill-McKee permutation of the resident CSR graph
int fcp_rcm_device(fcp_ctx *ctx, int pp_iters, int *new_of_old_dev);
int fcp_sort_pairs_int(fcp_ctx *ctx, const int *keys_in, int *keys_out,
                       const int *vals_in, int *vals_out, int n, int max_key);
