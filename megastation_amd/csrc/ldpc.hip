// LDPC consumer of the detection path's LLRs: layered normalized min-sum
// decoding of 5G NR code blocks (base graphs of 3GPP TS 38.212, ldpc_tables.h)
// behind the reference's LDPCDecoder contract (src/operators/ldpc/decoder.h:
// 73-86): per code block (Kb + mb) * Zc LLRs, binary16 or float, the first
// 2 * Zc (the punctured systematic bits) zero, LLR > 0 meaning bit 0; out the
// hard decisions of the Kb * Zc information bits, 32 per word, LSB first.
//
// One workgroup per code block, one thread per check node z < Zc of the
// current block row:
//   * the a-posteriori LLRs L of all (Kb + mb) * Zc variables live in LDS as
//     FP32; the circulant shift of every block comes reduced modulo Zc in the
//     kernel arguments, so it is a scalar load at a constant offset;
//   * the block rows are unrolled at compile time from the base graph's
//     structure (which columns a row touches), so the compressed check
//     messages of every row -- alpha*min1, alpha*min2, the index of the
//     minimum and the sign bits -- stay in registers across iterations;
//   * rows are processed in order (layered schedule) with one barrier
//     between rows: within a row the Zc checks touch disjoint variables;
//   * a row's LDS reads are all issued before the first is used, and the
//     search for the two smallest magnitudes is branch-free selects, so a row
//     costs one LDS round trip and no divergent branches.
// Per check and row, exactly the operations of the CPU oracle in the same
// order (round-to-nearest FP32 subtract, multiply and add, no contraction),
// so the decoded bits and the final LLRs equal the oracle's.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "eq_core.h"
#include "internal.h"
#include "ldpc_tables.h"

namespace mega {
namespace detail {
namespace {

// Row structure of a base graph: rows, information columns and the edge
// range of each row; the columns and shifts of the edges travel in
// LdpcParams::edge.
template <int BG>
struct Bg;
template <>
struct Bg<1> {
  static constexpr int rows = ldpc::kBg1Rows, kb = 22;
  static constexpr int start(int r) { return ldpc::kBg1RowStart[r]; }
};
template <>
struct Bg<2> {
  static constexpr int rows = ldpc::kBg2Rows, kb = 10;
  static constexpr int start(int r) { return ldpc::kBg2RowStart[r]; }
};

constexpr int kMaxEdges = 320;  // >= the edge count of either base graph

struct LdpcParams {
  const void* llr;
  uint64_t llr_stride;  // elements between code blocks
  uint32_t* out;
  uint64_t out_stride;  // words between code blocks
  uint32_t zc, mb, iters, half;
  float alpha;
  // every block of the base graph: the LDS byte offset of its column,
  // 4 * column * zc < 2^23, above its shift of the lifting-size set mod zc
  // (< 2^9)
  uint32_t edge[kMaxEdges];
};
static_assert(ldpc::kBg1Edges <= kMaxEdges && ldpc::kBg2Edges <= kMaxEdges, "edge table");
// The kernel's one argument, LdpcParams, starts the kernarg segment.
using KEdge = const __attribute__((address_space(4))) uint32_t*;
static_assert(offsetof(LdpcParams, edge) % 4 == 0 && sizeof(LdpcParams) <= 4096, "kernarg");

template <int BG>
__global__ __launch_bounds__(384) void ldpc_decode(const LdpcParams p) {
  using G = Bg<BG>;
  extern __shared__ float L[];  // (kb + mb) * zc LLRs, then the bit words
  const uint32_t zc = p.zc, mb = p.mb;
  const uint32_t nv = (G::kb + mb) * zc;
  uint32_t* bits = reinterpret_cast<uint32_t*>(L + nv);
  char* const Lb = reinterpret_cast<char*>(L);
  const uint32_t nbw = (G::kb * zc + 31) / 32;
  const uint32_t cw = blockIdx.x;
  const uint32_t z = threadIdx.x;
  const bool act = z < zc;

  for (uint32_t v = threadIdx.x; v < nv; v += blockDim.x) {
    const uint64_t i = cw * p.llr_stride + v;
    L[v] = p.half ? (float)static_cast<const _Float16*>(p.llr)[i]
                  : static_cast<const float*>(p.llr)[i];
  }
  for (uint32_t w = threadIdx.x; w < nbw; w += blockDim.x) bits[w] = 0u;
  __syncthreads();

  // compressed messages of this thread's check in every row
  float m1[G::rows], m2[G::rows];
  uint32_t pk[G::rows];  // sign bits 0..23, index of the minimum in bits 24..31
#pragma unroll
  for (int r = 0; r < G::rows; ++r) {
    m1[r] = 0.f;
    m2[r] = 0.f;
    pk[r] = 0u;
  }

  for (uint32_t it = 0; it < p.iters; ++it) {
    static_for<0, G::rows>([&](auto rc) {
      constexpr int r = decltype(rc)::value;
      constexpr int e0 = G::start(r), deg = G::start(r + 1) - G::start(r);
      if ((uint32_t)r >= mb) return;  // uniform
      if (act) {
        // the row's variables: the circulant maps check z to (z + shift) mod zc.
        // The edge words are scalar loads from the kernel arguments; they and
        // the addresses do not change between iterations, and a kernarg
        // address the compiler cannot see through keeps it from holding every
        // row's in registers.
        uint64_t ka = reinterpret_cast<uint64_t>(__builtin_amdgcn_kernarg_segment_ptr());
        __asm__ __volatile__("" : "+s"(ka));
        const KEdge ew = reinterpret_cast<KEdge>(ka + offsetof(LdpcParams, edge));
        uint32_t vi[deg];  // LDS byte addresses
        float lv[deg];
#pragma unroll
        for (int e = 0; e < deg; ++e) {
          const uint32_t w = ew[e0 + e];
          const uint32_t t = z + (w & 511u);  // < 2 zc
          vi[e] = (w >> 9) + 4 * min(t, t - zc);
        }
#pragma unroll
        for (int e = 0; e < deg; ++e) lv[e] = *reinterpret_cast<const float*>(Lb + vi[e]);
        // R_old of edge e: alpha*min2 at the old minimum's edge, alpha*min1
        // elsewhere (both >= +0), negated where bit e of the sign word is set;
        // the sign goes in as a bit, which equals the negation
        float q[deg];
        float min1 = __builtin_inff(), min2 = __builtin_inff();
        uint32_t im = 0, sb = 0u;
        const uint32_t amin = pk[r] >> 24;
#pragma unroll
        for (int e = 0; e < deg; ++e) {
          const float mag = (uint32_t)e == amin ? m2[r] : m1[r];
          const float rold = __uint_as_float(((pk[r] << (31 - e)) & 0x80000000u) |
                                             __float_as_uint(mag));
          const float qe = __fsub_rn(lv[e], rold);
          const float a = fabsf(qe);
          // a < min1: (min1, min2) <- (a, min1); else a < min2: min2 <- a.
          // With min1 <= min2 the new min2 is the median of the three.
          im = a < min1 ? (uint32_t)e : im;
          min2 = __builtin_amdgcn_fmed3f(min1, min2, a);
          min1 = fminf(min1, a);
          sb |= (qe < 0.f ? 1u : 0u) << e;
          q[e] = qe;
        }
        // new signs: each Q's sign XOR the XOR of all of them
        constexpr uint32_t all = (1u << deg) - 1u;
        const uint32_t nsg = __builtin_popcount(sb) & 1u ? sb ^ all : sb;
        const float n1 = __fmul_rn(p.alpha, min1), n2 = __fmul_rn(p.alpha, min2);
#pragma unroll
        for (int e = 0; e < deg; ++e) {
          const float mag = (uint32_t)e == im ? n2 : n1;
          const float rnew = __uint_as_float(((nsg << (31 - e)) & 0x80000000u) |
                                             __float_as_uint(mag));
          *reinterpret_cast<float*>(Lb + vi[e]) = __fadd_rn(q[e], rnew);
        }
        m1[r] = n1;
        m2[r] = n2;
        pk[r] = nsg | (im << 24);
      }
      __syncthreads();
    });
  }

  // hard decisions of the information bits, packed through LDS
  for (uint32_t k = threadIdx.x; k < (uint32_t)G::kb * zc; k += blockDim.x)
    if (L[k] < 0.f) atomicOr(&bits[k / 32], 1u << (k % 32));
  __syncthreads();
  for (uint32_t w = threadIdx.x; w < nbw; w += blockDim.x)
    p.out[cw * p.out_stride + w] = bits[w];
}

}  // namespace

// Lifting-size set of Zc (Table 5.3.2-1), or -1.
int ldpc_set_index(uint32_t zc) {
  static const uint32_t a[8] = {2, 3, 5, 7, 9, 11, 13, 15};
  for (int i = 0; i < 8; ++i)
    for (uint32_t z = a[i]; z <= 384; z *= 2)
      if (z == zc) return i;
  return -1;
}

hipError_t launch_ldpc_decode(const LdpcArgs& a, hipStream_t st, const char** why) {
  if (a.num_cb == 0) return hipSuccess;
  const int set = ldpc_set_index(a.zc);
  if (set < 0) {
    *why = "lifting size is not one of TS 38.212 Table 5.3.2-1";
    return hipErrorInvalidValue;
  }
  LdpcParams p;
  p.llr = a.llr;
  p.llr_stride = a.llr_stride;
  p.out = a.out;
  p.out_stride = a.out_stride;
  p.zc = a.zc;
  p.mb = a.mb;
  p.iters = a.iters;
  p.half = a.llr_bits == 16;
  p.alpha = a.alpha;
  const uint32_t kb = a.bg == 1 ? 22u : 10u;
  const uint32_t edges = a.bg == 1 ? ldpc::kBg1Edges : ldpc::kBg2Edges;
  const int16_t* row = a.bg == 1 ? ldpc::kBg1Shift[set] : ldpc::kBg2Shift[set];
  const int8_t* col = a.bg == 1 ? ldpc::kBg1Col : ldpc::kBg2Col;
  for (uint32_t e = 0; e < (uint32_t)kMaxEdges; ++e)
    p.edge[e] = e < edges ? (4u * col[e] * a.zc) << 9 | (uint32_t)(row[e] % (int)a.zc) : 0u;
  hipError_t e = hipSuccess;
  const size_t lds = (size_t)(kb + a.mb) * a.zc * sizeof(float) +
                     (size_t)(kb * a.zc + 31) / 32 * sizeof(uint32_t);
  const dim3 grid(a.num_cb), block((a.zc + 63) / 64 * 64);
  const void* fn = a.bg == 1 ? reinterpret_cast<const void*>(&ldpc_decode<1>)
                             : reinterpret_cast<const void*>(&ldpc_decode<2>);
  if (lds > 65536) {
    e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  if (a.bg == 1)
    hipLaunchKernelGGL(ldpc_decode<1>, grid, block, lds, st, p);
  else
    hipLaunchKernelGGL(ldpc_decode<2>, grid, block, lds, st, p);
  return hipGetLastError();
}

}  // namespace detail
}  // namespace mega
