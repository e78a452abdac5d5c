// Per-grid-size launch geometry shared by the forward and adjoint kernels.
#pragma once

namespace cn {

template <int N>
struct Cfg {
  // Row transforms: an N/2-point complex FFT per row, E values per thread, ROWS rows per block.
  static constexpr int L_ROW = N / 2;
  static constexpr int E_ROW = L_ROW >= 16 ? 16 : L_ROW;
  static constexpr int T_ROW = L_ROW / E_ROW;
  static constexpr int ROWS = (256 / T_ROW) < N ? (256 / T_ROW) : N;
  static constexpr int ROW_THREADS = ROWS * T_ROW;
  static constexpr int ROW_BLOCKS = N / ROWS;  // blocks per plane
  // Column transforms: an N-point complex FFT per column, E values per thread.
  static constexpr int E_COL = N >= 16 ? 16 : N;
  static constexpr int T_COL = N / E_COL;
  // default column tile width (complex columns per block of the column passes)
  static constexpr int TC_DEF = N >= 2048 ? 4 : (N >= 1024 ? 8 : 16);
  static constexpr int TC = TC_DEF < N / 2 ? TC_DEF : N / 2;
};

// Column tile widths a launcher may pick for size N (LDS: N*TC*8 + N*8 bytes <= 160 KiB).
template <int N, int TC>
constexpr bool col_tc_ok() {
  return TC <= N / 2 && (long)N * TC * 8 + (long)N * 8 <= 160 * 1024;
}

// XCD-aware tile order: workgroups are dealt round-robin to the 8 XCDs (separate L2s); give
// each XCD a contiguous run of tiles so that neighbouring tiles, which share 128-byte lines of
// every row, are fetched through the same L2.
template <int NTILES>
__device__ __forceinline__ int xcd_tile(int bx) {
  return (NTILES % 8 == 0) ? (bx % 8) * (NTILES / 8) + bx / 8 : bx;
}

}  // namespace cn
