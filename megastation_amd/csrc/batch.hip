// Batching form: ZF + equalize + demod of many cells of uneven shapes (each
// cell: its own U, B, N, sg, S, Q and LLR layout) in one hipGraph launch.
//
// Cells are grouped into classes that share a kernel configuration:
//   ZF classes  - users rounded up to 16-row tiles (UT = 1, 2, 4); one
//                 workgroup per (cell, group), the cell found from a prefix
//                 table of group counts (zf_core.h does the group's work);
//   EQ classes  - antenna count and the tile plan the autotuner picked for
//                 it (streamed W with MT 8-user M tiles per wave, or W held
//                 in registers); a persistent grid whose waves take units
//                 (cell, group, column tile, user group) from a work counter
//                 until it runs out, longest cells first.
// The modulation order is a per-cell runtime value (the demod epilogue
// switches on it), so it does not split classes.
#include "batch.h"
#include "zf_core.h"

namespace mega {
namespace detail {
namespace {

// Index of the cell whose range [prefix[c], prefix[c+1]) holds i (prefix has
// n+1 entries, n small): the same for the whole workgroup or wave.
__device__ __forceinline__ uint32_t find_cell(const uint32_t* prefix, uint32_t n, uint32_t i) {
  uint32_t lo = 0, hi = n;  // prefix[lo] <= i < prefix[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (prefix[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Threads per group, as mega_zf's: one wave up to 16 users, four up to 32,
// 16 for 33 to 64 (the same W bits on any of these counts).
template <int UT>
constexpr int cells_threads() { return UT == 1 ? 64 : UT == 2 ? 256 : 1024; }

// 17-32 users: three four-wave groups a CU, as in zf_mfma
template <int UT, bool A_LDS>
__global__ __launch_bounds__(cells_threads<UT>(), UT == 2 ? 3 : 1) void zf_cells(const BatchCell* cells,
                                                                const uint32_t* prefix,
                                                                uint32_t ncells) {
  constexpr int NTH = cells_threads<UT>();
  __shared__ ZfShared<UT> sh;
  extern __shared__ float sDyn[];
  const uint32_t c = find_cell(prefix, ncells, blockIdx.x);
  const BatchCell& cell = cells[c];
  const uint32_t g = blockIdx.x - prefix[c];
  const uint32_t U = cell.eq.U, B = cell.eq.B;
  float2* sA = reinterpret_cast<float2*>(sDyn);
  float2* part = reinterpret_cast<float2*>(sDyn + (A_LDS ? 2 * (U + 1) * B : 0));
  const uint64_t off = (uint64_t)g * U * B;
  const int fail = zf_group<UT, NTH, A_LDS>(cell.a + off, U, B, sh, sA, part, cell.w + off, U,
                                            nullptr);
  if (threadIdx.x == 0 && cell.info) cell.info[g] = fail;
}

// Equalize unit u of a cell (streamed W): u = ((g * col_tiles) + t) * nmg + mg.
template <int MT, int KC, int R, int NCH>
__device__ __forceinline__ void eq_cell_unit(const BatchCell& cell, uint32_t u) {
  const EqParams& p = cell.eq;
  const uint32_t mg = u % p.nmg;
  const uint32_t t = (u / p.nmg) % p.col_tiles;
  const uint32_t g = u / p.nmg / p.col_tiles;
  if (16 * t >= min(p.sg, p.N - g * p.sg) * p.S) return;  // past a partial group
  eq_unit_stream<MT, KC, R, NCH, 0>(p, 0, g, t, mg);  // runtime-Q epilogue
}

// Persistent streamed-W equalize over a class. The units are split into 8
// queues, one per XCD: queue x holds the x-th eighth of every cell's units
// (cells longest first), so the tiles of a group stay on one XCD and share
// its L2 copy of W_g. A wave serves the queue of its XCD (workgroups are
// dealt to XCDs round-robin, so blockIdx % 8) and then helps the others; it
// takes one unit per vector atomic on the queue's counter (lane 0, then
// broadcast) and exits when every queue is past its end.
template <int MT, int NCH>
__global__ __launch_bounds__(256) void eq_cells_stream(const BatchCell* cells,
                                                       const uint32_t* xprefix,
                                                       const uint32_t* xbase, uint32_t ncells,
                                                       uint32_t* counters) {
  const uint32_t home = blockIdx.x % 8;
  for (uint32_t k = 0; k < 8; ++k) {
    const uint32_t x = (home + k) % 8;
    const uint32_t* px = xprefix + x * (ncells + 1);
    const uint32_t end = px[ncells];
    for (;;) {
      uint32_t i = 0;
      if ((threadIdx.x & 63) == 0) i = atomicAdd(counters + x, 1u);
      i = __shfl(i, 0, 64);
      if (i >= end) break;
      const uint32_t c = find_cell(px, ncells, i);
      eq_cell_unit<MT, 2, 4, NCH>(cells[c], xbase[x * ncells + c] + (i - px[c]));
    }
  }
}

// Register form for narrow cells (all antennas in one wave's registers,
// nwk = 1): one wave per (cell, group, chunk of column tiles, user group).
template <int MT, int KS>
__global__ __launch_bounds__(256) void eq_cells_reg(const BatchCell* cells,
                                                    const uint32_t* prefix, uint32_t ncells) {
  const uint32_t unit = xcd_linear(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
  const uint32_t total = prefix[ncells];
  if (unit >= total) return;
  const uint32_t c = find_cell(prefix, ncells, unit);
  const BatchCell& cell = cells[c];
  const EqParams& p = cell.eq;
  const uint32_t u = unit - prefix[c];
  const uint32_t mg = u % p.nmg;
  const uint32_t chunk = (u / p.nmg) % p.chunks;
  const uint32_t g = u / p.nmg / p.chunks;
  const uint32_t nk = min(p.sg, p.N - g * p.sg);
  const uint32_t ntiles = (nk * p.S + 15) / 16;
  const uint32_t t0 = chunk * p.tiles_per_wg;
  const uint32_t t1 = min(t0 + p.tiles_per_wg, ntiles);
  if (t0 >= t1) return;
  const float2* W = p.csi + (uint64_t)g * p.U * p.B;
  float2 ynext[KS];
  eq_load_tile<KS>(p, 0, g, 0, t0, ynext);
  WRegs<MT, KS> ws;
  ws.load(W, p.U, p.U, p.B, mg, 0);
  eq_tiles<MT, KS, 0>(p, ws, 0, g, mg, 0, t0, t1, nullptr, ynext);  // runtime Q
}

__global__ void reset_counters(uint32_t* c, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c[i] = 0;
}

// A is staged in LDS for one-wave groups only (the class planner's choice).
template <int UT>
hipError_t prepare_zf_ut(const ZfClassLaunch& L) {
  if (L.lds <= 65536) return hipSuccess;
  const void* fn = reinterpret_cast<const void*>(&zf_cells<UT, false>);
  if constexpr (UT == 1)
    if (L.a_lds) fn = reinterpret_cast<const void*>(&zf_cells<UT, true>);
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.lds);
}

template <int UT>
hipError_t launch_zf_ut(const ZfClassLaunch& L, hipStream_t st) {
  const dim3 grid(L.blocks), block(cells_threads<UT>());
  if constexpr (UT == 1) {
    if (L.a_lds) {
      hipLaunchKernelGGL((zf_cells<UT, true>), grid, block, L.lds, st, L.cells, L.prefix,
                         L.ncells);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL((zf_cells<UT, false>), grid, block, L.lds, st, L.cells, L.prefix, L.ncells);
  return hipGetLastError();
}

template <int MT>
hipError_t launch_stream_mt(const EqClassLaunch& L, hipStream_t st) {
  const dim3 grid(L.blocks), block(256);
#define MEGA_EQ_CELLS(NCH)                                                                  \
  hipLaunchKernelGGL((eq_cells_stream<MT, NCH>), grid, block, 0, st, L.cells, L.xprefix, \
                     L.xbase, L.ncells, L.counters)
  switch (L.nch) {
    case 8: MEGA_EQ_CELLS(8); break;
    case 16: MEGA_EQ_CELLS(16); break;
    case 32: MEGA_EQ_CELLS(32); break;
    default: MEGA_EQ_CELLS(0); break;
  }
#undef MEGA_EQ_CELLS
  return hipGetLastError();
}

}  // namespace

hipError_t prepare_zf_class(const ZfClassLaunch& L) {
  switch (L.UT) {
    case 1: return prepare_zf_ut<1>(L);
    case 2: return prepare_zf_ut<2>(L);
    case 4: return prepare_zf_ut<4>(L);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_zf_class(const ZfClassLaunch& L, hipStream_t st) {
  if (L.blocks == 0) return hipSuccess;
  switch (L.UT) {
    case 1: return launch_zf_ut<1>(L, st);
    case 2: return launch_zf_ut<2>(L, st);
    case 4: return launch_zf_ut<4>(L, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_eq_class(const EqClassLaunch& L, hipStream_t st) {
  if (L.blocks == 0) return hipSuccess;
  if (L.kind == EqKind::Stream) {
    switch (L.MT) {
      case 1: return launch_stream_mt<1>(L, st);
      case 2: return launch_stream_mt<2>(L, st);
      case 4: return launch_stream_mt<4>(L, st);
      default: return hipErrorInvalidValue;
    }
  }
  const dim3 grid(L.blocks), block(256);
  switch (L.MT * 100 + L.KS) {
    case 108: hipLaunchKernelGGL((eq_cells_reg<1, 8>), grid, block, 0, st, L.cells, L.prefix, L.ncells); break;
    case 116: hipLaunchKernelGGL((eq_cells_reg<1, 16>), grid, block, 0, st, L.cells, L.prefix, L.ncells); break;
    case 208: hipLaunchKernelGGL((eq_cells_reg<2, 8>), grid, block, 0, st, L.cells, L.prefix, L.ncells); break;
    case 216: hipLaunchKernelGGL((eq_cells_reg<2, 16>), grid, block, 0, st, L.cells, L.prefix, L.ncells); break;
    case 408: hipLaunchKernelGGL((eq_cells_reg<4, 8>), grid, block, 0, st, L.cells, L.prefix, L.ncells); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_reset_counters(uint32_t* counters, uint32_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(reset_counters, dim3((n + 63) / 64), dim3(64), 0, st, counters, n);
  return hipGetLastError();
}

size_t zf_class_lds(uint32_t UT, uint32_t U, uint32_t B, bool a_lds) {
  switch (UT) {
    case 1: return zf_dyn_lds<1, 64>(U, B, a_lds);
    case 2: return zf_dyn_lds<2, 256>(U, B, a_lds);
    default: return zf_dyn_lds<4, cells_threads<4>()>(U, B, a_lds);
  }
}

size_t zf_class_static_lds(uint32_t UT) {
  switch (UT) {
    case 1: return sizeof(ZfShared<1>);
    case 2: return sizeof(ZfShared<2>);
    default: return sizeof(ZfShared<4>);
  }
}

}  // namespace detail
}  // namespace mega
