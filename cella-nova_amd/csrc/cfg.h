// Per-grid-size launch geometry shared by the forward and adjoint kernels.
#pragma once

namespace cn {

// Row transforms: an N/2-point complex FFT per row, E values per thread, ROWS rows per block.
template <int N, int E_REQ>
struct RowCfg {
  static constexpr int L = N / 2;
  static constexpr int E = E_REQ < L ? E_REQ : L;
  static constexpr int T = L / E;
  static constexpr int ROWS = T >= 256 ? 1 : ((256 / T) < N ? (256 / T) : N);
  static constexpr int THREADS = ROWS * T;
  static constexpr int BLOCKS = N / ROWS;  // blocks per plane
};

// Column transforms: an N-point complex FFT per column, E values per thread.
template <int N, int E_REQ>
struct ColCfg {
  static constexpr int E = E_REQ < N ? E_REQ : N;
  static constexpr int T = N / E;
};

// Built-in launch geometry (used when no tuned plan is given).
constexpr int default_col_tc(int N) {
  const int tc = N >= 2048 ? 4 : (N >= 1024 ? 8 : 16);
  return tc < N / 2 ? tc : N / 2;
}
template <int N>
struct DefaultGeom {
  static constexpr int FWD_E = 8;
  static constexpr int INV_E = 8;
  static constexpr int COL_E = 16;
  static constexpr int DIFF_TH = 16;
  static constexpr int TC = default_col_tc(N);
};

// Column geometries a launcher may pick for size N: at most 1024 threads per block and the
// padded exchange buffer (N + N/16) * TC * 8 bytes within 152 KiB of LDS (the rest of the
// 160 KiB holds the column kernels' twiddle tables, at most 8 KiB for N <= 2048).
constexpr long kColLdsBudget = 152 * 1024;
template <int N, int TC, int EC>
constexpr bool col_tc_ok() {
  return TC <= N / 2 && TC * ColCfg<N, EC>::T <= 1024 && (long)(N + N / 16) * TC * 8 <= kColLdsBudget;
}

// XCD-aware tile order: workgroups are dealt round-robin to the 8 XCDs (separate L2s); give
// each XCD a contiguous run of tiles so that neighbouring tiles, which share 128-byte lines of
// every row, are fetched through the same L2.
template <int NTILES>
__device__ __forceinline__ int xcd_tile(int bx) {
  return (NTILES % 8 == 0) ? (bx % 8) * (NTILES / 8) + bx / 8 : bx;
}

// Segmented column pass: columns per block (16 complex values are one 128-byte line per row;
// 32 at 512^2, where the column tiles are few: measured 2.5 % faster there, 1 % slower at
// 1024^2), the largest column-profile support radius it takes, and the segment length per grid
// size (0: full-length column transforms only).
constexpr int seg_tc(int N) { return N >= 1024 ? 16 : 32; }
// Values per thread of the segment transforms: 8 at 1024^2, where 16 leaves two-wavefront
// blocks too few to fill the CUs (column pass 17.6 -> 15.9 us); 16 elsewhere (8 was 2 % slower
// for the 512^2 batched search and 5 % slower at 2048^2).
constexpr int seg_e(int N) { return N == 1024 ? 8 : 16; }
constexpr int kSegRcap = 64;
constexpr int seg_m(int N) { return N == 2048 ? 256 : (N == 1024 || N == 512) ? 128 : 0; }
// The largest support radius a segment takes: half of its rows stay valid outputs.
constexpr int seg_rmax(int N) { return seg_m(N) / 4 < kSegRcap ? seg_m(N) / 4 : kSegRcap; }

}  // namespace cn
