// Kernel argument blocks and launchers of the ICON core (host + device view).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace icon {

constexpr int kTapsK = 14;     // Gaussian window taps per dimension
constexpr int kTapsKBK = 6;    // Kaiser-Bessel window taps per dimension
constexpr int kMaxSlotK = 3;

struct RowFwdArgs {
  int rows, N0, N1;
  int cplx;        // 0: real image m (with optional update), 1: complex rhat
  int update;      // 0 none, 1 INFR, 2 ICON (threshold)
  int data_type;
  float thr;
  float* m;
  const float2* rhat;   // cplx: the complex image
  const float* grad;    // update: Re(rhat) [S][N0][N1]
  const double* alpha;
  const float* ck;
  const float2* tw;
  float2* out;     // [S][N0][NR]
  int zlo, zhi;    // out columns [zlo, zhi) are not stored (no interpolation reads them)
};

struct RowAdjArgs {
  int rows, N0, N1;
  int mode;            // 0: write image (A^H result), 1: rhat = image - ahwb, row sums
  const float2* in;    // [S][N0][NR]
  const float* ck;
  const float2* tw;
  float2* img;         // mode 0 output
  const float2* ahwb;  // mode 1
  float* grad;         // mode 1 output: Re(rhat) [S][N0][N1], all the update reads
  double* rowpart;     // mode 1: sum |rhat|^2 per row
  float2* fwd_out;     // mode 1, or null: y-FFT of rhat/ck padded to NR, per row [S][N0][NR]
  int zlo, zhi;        // columns [zlo, zhi) of in are zero (not loaded) and of fwd_out not stored
};

struct ColArgs {
  int S, N0, N1, M, n1, E;
  int K;                   // window taps per dimension (kTapsK or kTapsKBK)
  const float2* in;        // fwd: tmp [S][N0][n1]
  float2* out;             // adj: tmp [S][N0][n1]
  const float2* tw;
  const int* ent_off;
  const int* tile_order;   // adj: [nblk] tile of each block group (blocks bid / S), largest first
  const int4* ent;
  const int* xoff;         // adj: [nblk][n0+1] row offsets into ent
  const float* phix;
  const float* phiy;
  const int* ix0;
  const float* ent_px;     // [E][K], interleaved per tile
  const float* ent_py;     // [E][C], interleaved per tile
  // fwd: the entries and their records in the interpolation's order
  const int4* fent;
  const float* fent_px;
  const float* fent_py;
  float2* part;            // fwd: [S][M][3]
  const float2* fw;        // adj: [S][E] weighted value per bin entry
  int zlo, zhi;            // adj: tiles inside columns [zlo, zhi) store nothing (their zeros are implied)
  int split_min;           // adj: tiles with at least this many entries split their rows' entries in two
};

struct CombineArgs {
  int S, M, N0, nz, E;
  int mode;              // 0: after A m, 1: after A rhat (step size), 2: initial b, 3: plain sum
  const float2* part;    // [S][M][3]
  const int8_t* nslot;
  const float* w;
  const float2* b;       // [S][M]
  float2* fw;            // mode 0/2: [S][E] weighted value per bin entry; mode 3: [S][M] sums
  const int* ent_of_node;  // [M][3] bin entries of each node
  double* bnorm;         // [S]
  double* resid;         // mode 0: trace base or null
  int resid_stride, resid_index;
  const double* rowpart; // mode 1: [S][N0]
  double* alpha;         // mode 1: [S]
  double* alpha_trace;   // mode 1: [S][iters] or null
  int alpha_stride, alpha_index;
  float2* frep;          // mode 0: [S][N0] unweighted values at the omit tilt, or null
  int iomit;
};

struct SinoArgs {
  int rows, N0;
  int K;                 // window taps
  int nz, skip;          // tilts per slice; input row left out (cross validation), else -1
  const float* sino;     // [S][nz][N0]
  const float* ck1;
  const float* phi1;
  const int* i10;
  const float2* tw;
  float2* b;             // [S][nz*N0]
};

struct ReprojArgs {
  int S, N0;
  int K;                 // window taps
  float sigma;
  const float2* frep;    // [S][N0]
  const float* ck1;
  const float* phi1;
  const int* i10;
  const float2* tw;
  float* out;            // [S][N0]
};

// Per-launch fields of a packed launch over several series (the argument
// blocks of the series are read from a device table).
struct RowFwdLaunch {
  int cplx, update, data_type;
  float thr;
};

struct CombineLaunch {
  int mode, resid_index, alpha_index;
  int with_resid, with_frep;   // 0: the series' resid / frep pointers are ignored
};

// While set (non-null) on the calling thread, the next launches of the
// iteration kernels (row_fwd, col, combine, row_adj, sino) go through
// hipExtLaunchKernel with these events, which then hold the kernel's own start
// and end on the device clock. Used only for timing, never under capture.
void set_launch_events(hipEvent_t start, hipEvent_t stop);

hipError_t launch_row_fwd(int NR, const RowFwdArgs& a, hipStream_t st);
hipError_t launch_row_adj(int NR, const RowAdjArgs& a, hipStream_t st);
hipError_t launch_col(int NC, int C, bool adjoint, const ColArgs& a, hipStream_t st);
hipError_t launch_combine(const CombineArgs& a, hipStream_t st);
hipError_t launch_sino(int NR, const SinoArgs& a, hipStream_t st);
hipError_t launch_reproj(int NR, const ReprojArgs& a, hipStream_t st);
hipError_t launch_fft_rows(int NR, float2* data, int rows, int sign, const float2* tw, hipStream_t st);
hipError_t launch_fft_cols(int NC, int C, float2* data, int n1, int count, int sign, const float2* tw,
                           hipStream_t st);
hipError_t launch_zero(float* p, size_t n, hipStream_t st);
// fw_ent[s][e] = f[s][node of e]
hipError_t launch_node_to_ent(const float2* f, float2* fw_ent, const int4* ent, int S, int M, int E, hipStream_t st);
hipError_t launch_copy(float* dst, const float* src, size_t n, hipStream_t st);
hipError_t launch_copy4(float4* dst, const float4* src, size_t n, hipStream_t st);
// Packed launches: `n` blocks, block b running block items[b].y of the series
// descs[items[b].x] (both tables in device memory).  Rows per block of the
// row kernels (the unit of their items) by row length:
int row_fwd_rows_per_block(int NR);
int row_adj_rows_per_block(int NR);
int sino_rows_per_block(int NR);
hipError_t launch_row_fwd_multi(int NR, const RowFwdArgs* descs, const int2* items, int n, const RowFwdLaunch& L,
                                hipStream_t st);
hipError_t launch_row_adj_multi(int NR, const RowAdjArgs* descs, const int2* items, int n, int mode, bool fwd,
                                hipStream_t st);
hipError_t launch_col_multi(int NC, int C, int K, bool adjoint, const ColArgs* descs, const int2* items, int n,
                            hipStream_t st);
hipError_t launch_combine_multi(const CombineArgs* descs, const int2* items, int n, const CombineLaunch& L,
                                hipStream_t st);
hipError_t launch_sino_multi(int NR, const SinoArgs* descs, const int2* items, int n, hipStream_t st);
hipError_t launch_reproj_multi(int NR, const ReprojArgs* descs, const int2* items, int n, hipStream_t st);
bool col_config_ok(int NC, int C, int K);
// A column tile of C grid columns usable on an n0 x n1 grid with K taps.
bool col_tile_ok(int NC, int n1, int C, int K);
hipError_t set_col_attrs(int NC, int C, int K);
hipError_t set_col_multi_attrs(int NC, int C, int K);

}  // namespace icon
