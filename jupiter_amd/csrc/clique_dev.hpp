// clique_dev.hpp — device routines of the clique engine (bitmap build,
// clique recursion), shared by the single-query kernels and the batched /
// serving kernel.  See clique_kernels.hip for the design.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

#include "kernels.hpp"

namespace jk {

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
  uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
  lo = __shfl_xor(lo, m);
  hi = __shfl_xor(hi, m);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += shfl_xor_u64(v, m);
  return v;
}

__device__ __forceinline__ int wave_incl_scan(int x) {
  const int l = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o);
    if (l >= o) x += y;
  }
  return x;
}

// last index i in [0, n) with a[i] <= f, given a[0] <= f < a[n]
__device__ __forceinline__ int seg_search(const int* a, int n, int f) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (a[mid] <= f) lo = mid; else hi = mid;
  }
  return lo;
}

// Cursor form of seg_search for increasing f: a[i] <= f on entry; returns
// the last j >= i with a[j] <= f: two single steps forward (most moves), then
// a galloping search, so a jump over m segments costs O(log m) reads.
__device__ __forceinline__ int seg_advance(const int* a, int n, int i, int f) {
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    if (a[i + 1] > f) return i;
    ++i;
  }
  int lo = i, step = 1;
  while (lo + step < n && a[lo + step] <= f) {
    lo += step;
    step <<= 1;
  }
  int hi = min(n, lo + step);  // a[hi] > f, or hi == n
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] <= f) lo = mid; else hi = mid;
  }
  return lo;
}

// first index p in [0, n) with a[p] >= y (n if none)
template <class T>
__device__ __forceinline__ int lower_bound_dev(const T* a, int n, int y) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] < y) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ bool bsearch_global(const int32_t* __restrict__ a, int64_t n, int32_t y) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    int32_t v = a[mid];
    if (v < y) lo = mid + 1; else hi = mid;
  }
  return lo < n && a[lo] == y;
}

// Group of G waves sharing one root (G == 1: a lone wave; G > 1: the block).
template <int G>
struct Grp {
  static constexpr int size = 64 * G;
  __device__ static int rank() { return G == 1 ? lane_id() : (int)threadIdx.x; }
  __device__ static int wave() { return G == 1 ? 0 : (int)(threadIdx.x >> 6); }
  __device__ static void sync() {
    if constexpr (G == 1) __syncwarp(); else __syncthreads();
  }
};

// Exclusive scan of a[0..n) in place (n <= 64 * G * per-thread chunk), sets a[n] = total.
template <int G>
__device__ void group_excl_scan(int* a, int n, int* wtot) {
  const int r = Grp<G>::rank();
  const int per = (n + Grp<G>::size - 1) / Grp<G>::size;
  const int b = r * per;
  const int e = min(n, b + per);
  int s = 0;
  for (int t = b; t < e; ++t) s += a[t];
  int incl = wave_incl_scan(s);
  int base = 0;
  if constexpr (G > 1) {
    const int w = Grp<G>::wave();
    if (lane_id() == 63) wtot[w] = incl;
    __syncthreads();
    for (int k = 0; k < w; ++k) base += wtot[k];
    if (w == G - 1 && lane_id() == 63) wtot[G] = base + incl;
  }
  int run = base + incl - s;
  for (int t = b; t < e; ++t) {
    int v = a[t];
    a[t] = run;
    run += v;
  }
  if constexpr (G == 1) {
    if (lane_id() == 63) a[n] = incl;
  }
  Grp<G>::sync();
  if constexpr (G > 1) {
    if (threadIdx.x == 0) a[n] = wtot[G];
    __syncthreads();
  }
}

// Count (depth)-cliques-with-prefix over single-word rows `R`: depth 1 is
// popcount(C); deeper levels branch on every member x of C with C & R[x].
// With BYTES the weights w[x] of every visited member are summed instead
// (work-model bytes of the reference nest).
template <int DEPTH, bool BYTES, class IDX = int>
__device__ uint64_t rec1(uint64_t C, const uint64_t* R, const uint32_t* w, const IDX* map = nullptr) {
  if constexpr (DEPTH <= 0) {
    return 0;
  } else if constexpr (DEPTH == 1 && !BYTES) {
    return (uint64_t)__popcll(C);
  } else {
    uint64_t s = 0;
    if constexpr (!BYTES) {
      if (__popcll(C) < DEPTH) return 0;
    }
    // Positions follow vertex ids, not the DAG order, so a successor of x
    // may sit at a lower position: branch on C & R[x] with the whole of C.
    const uint64_t C0 = C;
    while (C) {
      int x = __builtin_ctzll(C);
      C &= C - 1;
      if constexpr (BYTES) s += w[map ? (int)map[x] : x];
      if constexpr (DEPTH > 1) s += rec1<DEPTH - 1, BYTES, IDX>(C0 & R[x], R, w, map);
    }
    return s;
  }
}

// Multi-word fallback: words of the current set are the AND of the rows of
// the chosen positions p[0..NP), recomputed on the fly (no register arrays).
template <int DEPTH, int NP, bool BYTES, class RowT>
__device__ uint64_t recW(const RowT& R, int wlo, int whi, const int* p, const uint32_t* wt) {
  uint64_t s = 0;
  for (int w = wlo; w <= whi; ++w) {
    uint64_t m = R(p[0], w);
#pragma unroll
    for (int k = 1; k < NP; ++k) m &= R(p[k], w);
    if constexpr (DEPTH == 1 && !BYTES) {
      s += (uint64_t)__popcll(m);
    } else {
      while (m) {
        int x = (w << 6) | __builtin_ctzll(m);
        m &= m - 1;
        if constexpr (BYTES) s += wt[x];
        if constexpr (DEPTH > 1) {
          int q[NP + 1];
#pragma unroll
          for (int k = 0; k < NP; ++k) q[k] = p[k];
          q[NP] = x;
          s += recW<DEPTH - 1, NP + 1, BYTES>(R, wlo, whi, q, wt);
        }
      }
    }
  }
  return s;
}

// Word w of bitmap row i sits at i * S + (w ^ ((i >> sh) & X)): S = W when
// W is odd or a power of two, else W + 1 (odd); for a power of two W the
// word is XOR-swizzled by the row bits above the 32 / W rows that one pass
// of the stride spreads over 32 eight-byte bank slots.  Either way lanes that
// read the same word of 32 consecutive rows hit 32 different slots.  d * S
// never exceeds 64 * WMAX * WMAX, the rows' allocation.
struct RowLayout {
  int S, X, sh;
  __device__ explicit RowLayout(int W) {  // W >= 0 (W == 0: a root without members)
    const bool p2 = W > 0 && (W & (W - 1)) == 0;
    S = ((W & 1) || p2) ? W : W + 1;
    X = p2 ? W - 1 : 0;
    sh = 0;
    while (sh < 5 && (W << sh) < 32) ++sh;
  }
  __device__ int operator()(int i, int w) const { return i * S + (w ^ ((i >> sh) & X)); }
};
// Row accessors: word w of row i, and its 32-bit half h (bits 32h .. 32h + 31).
// A cursor fixes row i and half h: its byte base and the row's word swizzle
// pre-scaled to bytes, so a read at word w costs one xor-add
// (8 (w ^ s) = 8w ^ 8s).
struct RowCur {
  const unsigned char* b;
  uint32_t sw;
};
struct RowsLds {
  uint64_t* p;
  RowLayout lay;
  __device__ uint64_t operator()(int i, int w) const { return p[lay(i, w)]; }
  __device__ uint32_t half(int i, int w, int h) const {
    return reinterpret_cast<const uint32_t*>(p)[2 * lay(i, w) + h];
  }
  __device__ RowCur cur(int i, int h) const {
    return {reinterpret_cast<const unsigned char*>(p + (size_t)i * lay.S) + 4 * h, 8u * (uint32_t)((i >> lay.sh) & lay.X)};
  }
  __device__ uint32_t at(const RowCur& c, int w) const {
    return *reinterpret_cast<const uint32_t*>(c.b + ((8u * (uint32_t)w) ^ c.sw));
  }
};
struct RowsGlobal {
  uint64_t* p;
  RowLayout lay;
  __device__ uint64_t operator()(int i, int w) const { return __builtin_nontemporal_load(p + lay(i, w)); }
  __device__ uint32_t half(int i, int w, int h) const {
    return __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(p) + 2 * lay(i, w) + h);
  }
  __device__ RowCur cur(int i, int h) const {
    return {reinterpret_cast<const unsigned char*>(p + (size_t)i * lay.S) + 4 * h, 8u * (uint32_t)((i >> lay.sh) & lay.X)};
  }
  __device__ uint32_t at(const RowCur& c, int w) const {
    return __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(c.b + ((8u * (uint32_t)w) ^ c.sw)));
  }
};

// Block reduction of a per-thread uint64 and one atomicAdd per block.
template <int NT>
__device__ void block_add(uint64_t v, uint64_t* out) {
  __shared__ uint64_t part[NT / 64];
  v = wave_sum_u64(v);
  if (lane_id() == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < NT / 64; ++k) s += part[k];
    if (s) atomicAdd((unsigned long long*)out, (unsigned long long)s);
  }
}

// Next set bit at or after `from` in the first CW words of c (-1 if none).
template <int CWMAX>
__device__ __forceinline__ int next_bit(const uint64_t (&c)[CWMAX], int CW, int from) {
#pragma unroll
  for (int w = 0; w < CWMAX; ++w) {
    if (w < CW && from < (w + 1) * 64) {
      uint64_t m = c[w];
      if (from > w * 64) m &= (~0ull) << (from - w * 64);
      if (m) return w * 64 + __builtin_ctzll(m);
    }
  }
  return -1;
}

// Clique recursion over compressed multi-word rows C (CW words per row):
// the candidate set lives in registers, rows in LDS.
template <int DEPTH, int CWMAX, bool BYTES>
__device__ uint64_t recC(const uint64_t (&c)[CWMAX], int CW, const uint64_t* C, const uint32_t* wt) {
  if constexpr (DEPTH <= 0) {
    return 0;
  } else if constexpr (DEPTH == 1 && !BYTES) {
    uint64_t s = 0;
#pragma unroll
    for (int w = 0; w < CWMAX; ++w)
      if (w < CW) s += (uint64_t)__popcll(c[w]);
    return s;
  } else {
    uint64_t s = 0;
    if constexpr (!BYTES) {
      int pc = 0;
#pragma unroll
      for (int w = 0; w < CWMAX; ++w)
        if (w < CW) pc += __popcll(c[w]);
      if (pc < DEPTH) return 0;
    }
    for (int x = next_bit<CWMAX>(c, CW, 0); x >= 0; x = next_bit<CWMAX>(c, CW, x + 1)) {
      if constexpr (BYTES) s += wt[x];
      if constexpr (DEPTH > 1) {
        uint64_t nx[CWMAX];
#pragma unroll
        for (int w = 0; w < CWMAX; ++w) nx[w] = (w < CW) ? (c[w] & C[x * CW + w]) : 0ull;
        s += recC<DEPTH - 1, CWMAX, BYTES>(nx, CW, C, wt);
      }
    }
    return s;
  }
}

// Clique recursion over the original multi-word rows restricted to the
// word window [wlo, wlo + nw) of the first row: the candidate set lives in
// WM registers, each branch reads nw words of one row.
template <int DEPTH, int WM, bool BYTES, class RowT>
__device__ uint64_t recR(const uint64_t (&c)[WM], int wlo, int nw, const RowT& R, const uint32_t* wt) {
  if constexpr (DEPTH <= 0) {
    return 0;
  } else if constexpr (DEPTH == 1 && !BYTES) {
    uint64_t s = 0;
#pragma unroll
    for (int w = 0; w < WM; ++w)
      if (w < nw) s += (uint64_t)__popcll(c[w]);
    return s;
  } else {
    uint64_t s = 0;
    if constexpr (!BYTES) {
      int pc = 0;
#pragma unroll
      for (int w = 0; w < WM; ++w)
        if (w < nw) pc += __popcll(c[w]);
      if (pc < DEPTH) return 0;
    }
    for (int b = next_bit<WM>(c, nw, 0); b >= 0; b = next_bit<WM>(c, nw, b + 1)) {
      const int x = wlo * 64 + b;
      if constexpr (BYTES) s += wt[x];
      if constexpr (DEPTH > 1) {
        uint64_t nx[WM];
        // rows hold successors only (positions after x): skip the words before
        const int ws = (b + 1) >> 6;
#pragma unroll
        for (int w = 0; w < WM; ++w) nx[w] = (w >= ws && w < nw) ? (c[w] & R(x, wlo + w)) : 0ull;
        s += recR<DEPTH - 1, WM, BYTES>(nx, wlo, nw, R, wt);
      }
    }
    return s;
  }
}

typedef int v8i32 __attribute__((ext_vector_type(8)));
typedef float v16f32 __attribute__((ext_vector_type(16)));

// The 5-clique count runs on the fp4 (e2m1) matrix cores: 0/1 operands are
// exact in e2m1 and the f32 accumulators hold the counts (at most 2^11 here)
// exactly.  v_mfma_scale_f32_32x32x64_f8f6f4 with fp4 operands covers 64
// positions per instruction from 16 operand bytes per lane: twice the
// positions per cycle of the int8 instructions and half their operand bytes.
//
// Operand encoding, in registers: bit e = 4n + d of a lane's 32 positions is
// element e of the operand (dword d, nibble n), so a set bit only has to be
// masked, or moved inside its nibble.  The A side keeps bits 0..2 of each
// nibble where they are (e2m1 0.5, 1.0, 2.0) and moves bit 3 to bit 0 (0.5);
// the B side holds the reciprocal at every element (2.0, 1.0, 0.5, 2.0), so
// each product of two set elements is exactly 1 and a product counts common
// positions.
__device__ __forceinline__ v8i32 enc_a(uint32_t x) {
  v8i32 v;
  v[0] = (int)(x & 0x11111111u);
  v[1] = (int)(x & 0x22222222u);
  v[2] = (int)(x & 0x44444444u);
  v[3] = (int)((x >> 3) & 0x11111111u);
  v[4] = v[5] = v[6] = v[7] = 0;  // the fp4 form reads the first four dwords
  return v;
}
__device__ __forceinline__ v8i32 enc_b(uint32_t x) {
  v8i32 v;
  v[0] = (int)((x << 2) & 0x44444444u);
  v[1] = (int)(x & 0x22222222u);
  v[2] = (int)((x >> 2) & 0x11111111u);
  v[3] = (int)((x >> 1) & 0x44444444u);
  v[4] = v[5] = v[6] = v[7] = 0;
  return v;
}

// D += A B^T on 32 x 32 x 64 fp4 operands with unit scales (E8M0 127 = 2^0).
__device__ __forceinline__ v16f32 mma_fp4(const v8i32& a, const v8i32& b, const v16f32& c) {
  return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, 4, 4, 0, 127, 0, 127);
}

// Matrix-core instructions run on the whole wave whatever EXEC holds, so
// every condition that guards one must be wave-uniform and scalar (a real
// branch), never a lane mask: values the compiler cannot prove uniform are
// passed through readfirstlane.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ v16f32 zero16() {
  v16f32 z;
#pragma unroll
  for (int k = 0; k < 16; ++k) z[k] = 0.f;
  return z;
}

// Sum of the accumulator entries [y][x] of block Y against column member x
// whose pair is a DAG edge x -> y (bit y of R[x]).  Lane (col, half) holds
// the entries of members ym = 32Y + 8g + 4half + e (g, e < 4): four runs of
// four consecutive J entries, read as 8-byte words.  Row x is read at the
// block's first two words; members further on (rare) read their own word.
// Entries are counts of at most 2^11 (row width 2048), so the f32 sum of 16
// is exact (checked by the static_assert in process_root).
template <class RowT>
__device__ __forceinline__ uint32_t edge_masked_sum(const RowT& R, const int16_t* J, int n, int whi, int Y, int x,
                                                    int half, const v16f32& acc) {
  float s = 0.f;
  if (x >= 0) {
    const int w0 = uni(J[32 * Y]) >> 6;
    const uint64_t r0 = R(x, w0), r1 = R(x, min(w0 + 1, whi));
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int base = 32 * Y + 8 * g + 4 * half;
      const uint64_t four = *reinterpret_cast<const uint64_t*>(J + base);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int yy = (int)(int16_t)(uint16_t)(four >> (16 * e));
        if (base + e < n) {
          const int w = yy >> 6;
          const uint64_t word = w == w0 ? r0 : (w == w0 + 1 ? r1 : R(x, w));
          if ((word >> (yy & 63)) & 1ull) s += acc[4 * g + e];
        }
      }
    }
  }
  return (uint32_t)s;
}

// A-side operand with a single set element at position y (when y lies in
// the 32 positions of word q's half `half`): E_Y in E_Y * R_X^T, whose entry
// [y][x] is bit y of R[x], i.e. the edge x -> y.
__device__ __forceinline__ v8i32 one_hot_fp4(int y, int q, int half) {
  const bool in = y >= 0 && (y >> 5) == 2 * q + half;
  return enc_a(in ? 1u << (y & 31) : 0u);
}

// Entries are counts of at most 2^11 and 0/1 masks: the f32 sum of 16 products
// is exact.
__device__ __forceinline__ uint32_t dot16(const v16f32& c, const v16f32& m) {
  float s = 0.f;
#pragma unroll
  for (int a = 0; a < 16; ++a) s += c[a] * m[a];
  return (uint32_t)s;
}

// Triangles of the DAG induced on S = R[i] (rows ranked in DAG order, so
// every row only has bits above its own position), on the matrix cores:
//   sum over member pairs x -> y of |R[y] ∩ R[i] ∩ R[x]|
// = sum over 32x32 blocks (X <= Y) of  mask(x -> y) . (A_Y B_X^T)[y][x],
// where A_Y holds the rows R[y] ∩ R[i] of the Y block's members and B_X the
// rows R[x] of the X block's members, over the row's words from the first
// that can hold a common successor to whi, one fp4 MFMA per 64-position word.
// Lane l supplies member 32B + l%32 and the 32 positions 32 * (l / 32) .. +31
// of the word; accumulator element a of lane l is entry
// [8 (a / 4) + 4 (l / 32) + a % 4][l % 32].  Absent members (past n) read row
// zr, which is empty (the last row of the DAG order), so every load and
// product is unconditional; the first word's product starts from a zero C
// operand, so the accumulators need no clearing.  J lists S's positions in
// ascending order (n > 0 entries); every lane of the wave must call this.
// This form keeps one accumulator per block pair and reads the edge mask from
// the rows, bit by bit (the 256-wide roots: rows of 4 words, where the
// members of a block mostly share a word and registers decide occupancy).
template <class RowT>
__device__ uint64_t tri_mfma1_bits(const RowT& R, const int16_t* J, int n_, int i_, int whi_, int zr_) {
  const int n = uni(n_), i = uni(i_), whi = uni(whi_), zr = uni(zr_);
  const int ln = lane_id();
  const int col = ln & 31, half = ln >> 5;
  const int T = (n + 31) >> 5;
  uint64_t s = 0;
  for (int Y = 0; Y < T; ++Y) {
    const int ay = 32 * Y + col;
    const int y = ay < n ? (int)J[ay] : -1;
    const int q0 = (uni(J[32 * Y]) + 1) >> 6;  // R[y] is empty up to the block's first member
    if (q0 > whi) continue;
    const int ry = y >= 0 ? y : zr;
    for (int X = 0; X <= Y; ++X) {
      const int bx = 32 * X + col;
      const int x = bx < n ? (int)J[bx] : -1;
      const int rx = x >= 0 ? x : zr;
      // word q + 1's row halves are loaded before word q's table reads
      // (two register sets alternate, so the loop moves no values)
      const RowCur ki = R.cur(i, half), ky = R.cur(ry, half), kx = R.cur(rx, half);
      uint32_t va[3] = {R.at(ki, q0), R.at(ky, q0), R.at(kx, q0)}, vb[3];
      auto step = [&](int q, const v16f32& c, const uint32_t(&v)[3], uint32_t(&nx)[3]) -> v16f32 {
        const uint32_t sa = v[1] & v[0], sb = v[2];
        const int wn = min(q + 1, whi);
        nx[0] = R.at(ki, wn);
        nx[1] = R.at(ky, wn);
        nx[2] = R.at(kx, wn);
        return mma_fp4(enc_a(sa), enc_b(sb), c);
      };
      v16f32 acc = step(q0, zero16(), va, vb);
      int q = q0 + 1;
      for (; q + 1 <= whi; q += 2) {
        acc = step(q, acc, vb, va);
        acc = step(q + 1, acc, va, vb);
      }
      if (q <= whi) acc = step(q, acc, vb, va);
      s += edge_masked_sum(R, J, n, whi, Y, x, half, acc);
    }
  }
  return s;
}

// Triangles of the DAG induced on S = R[i] (rows ranked in DAG order, so
// every row only has bits above its own position), on the matrix cores:
//   sum over member pairs x -> y of |R[y] ∩ R[i] ∩ R[x]|
// = sum over 32x32 blocks (X <= Y) of  mask(x -> y) . (A_Y B_X^T)[y][x],
// where A_Y holds the rows R[y] ∩ R[i] of the Y block's members and B_X the
// rows R[x] of the X block's members, over the row's words from the one
// holding the block's first member to whi, one fp4 MFMA per 64-position
// word.  The mask is the product E_Y B_X^T of rows one-hot at the Y members'
// own positions, accumulated on the words that hold them (qa .. qb) with the
// same B operands: its entry [y][x] is bit y of R[x], the edge x -> y.
// Lane l supplies member 32B + l%32 and the 32 positions 32 * (l / 32) .. +31
// of the word; accumulator element a of lane l is entry
// [8 (a / 4) + 4 (l / 32) + a % 4][l % 32].  Absent members (past n) read row
// zr, which is empty (the last row of the DAG order), so every load and
// product is unconditional; the first word's products start from a zero C
// operand, so the accumulators need no clearing.  J lists S's positions in
// ascending order (n > 0 entries); every lane of the wave must call this.
// This form keeps one block pair at a time (the 512- and 1024-wide roots,
// where registers decide occupancy); on rows of 512 positions it replaces
// sixteen bit tests per lane over up to eight words (tri_mfma1_bits) with one
// more product per word that holds a member of the block (P14 on R-MAT 22,
// roots [0, 2^20): 512-wide tier 21.5 -> 14.3 ms), and on the 1024-wide
// count its registers allow 12 waves per root where the 2x2 tiles of
// tri_mfma held 8 (23.0 -> 20.7 ms).
template <class RowT>
__device__ uint64_t tri_mfma1(const RowT& R, const int16_t* J, int n_, int i_, int whi_, int zr_) {
  const int n = uni(n_), i = uni(i_), whi = uni(whi_), zr = uni(zr_);
  const int ln = lane_id();
  const int col = ln & 31, half = ln >> 5;
  const int T = (n + 31) >> 5;
  uint64_t s = 0;
  for (int Y = 0; Y < T; ++Y) {
    const int ay = 32 * Y + col;
    const int y = ay < n ? (int)J[ay] : -1;
    // words qa .. qb hold the block's members (qb <= whi: S lies in row i)
    const int qa = uni(J[32 * Y]) >> 6, qb = uni(J[min(n - 1, 32 * Y + 31)]) >> 6;
    const int ry = y >= 0 ? y : zr;
    // the one-hot row of member y: bit y & 31 of half (y >> 5) & 1 of word y >> 6
    const uint32_t hbit = (y >= 0 && ((y >> 5) & 1) == half) ? 1u << (y & 31) : 0u;
    const int hq = y >= 0 ? y >> 6 : -1;
    for (int X = 0; X <= Y; ++X) {
      const int bx = 32 * X + col;
      const int rx = bx < n ? (int)J[bx] : zr;
      // word q + 1's row halves are loaded before word q's table reads
      // (two register sets alternate, so the loop moves no values)
      const RowCur ki = R.cur(i, half), ky = R.cur(ry, half), kx = R.cur(rx, half);
      uint32_t va[3] = {R.at(ki, qa), R.at(ky, qa), R.at(kx, qa)}, vb[3];
      v16f32 acc, msk;
      auto step = [&](int q, bool first, bool with_mask, const uint32_t(&v)[3], uint32_t(&nx)[3]) {
        const uint32_t sa = v[1] & v[0], sb = v[2];
        const int wn = min(q + 1, whi);
        nx[0] = R.at(ki, wn);
        nx[1] = R.at(ky, wn);
        nx[2] = R.at(kx, wn);
        const v8i32 b = enc_b(sb);
        const v16f32 z = zero16();
        acc = mma_fp4(enc_a(sa), b, first ? z : acc);
        if (with_mask) msk = mma_fp4(enc_a(hq == q ? hbit : 0u), b, first ? z : msk);
      };
      step(qa, true, true, va, vb);
      int q = qa + 1;
      for (; q + 1 <= whi; q += 2) {
        step(q, false, q <= qb, vb, va);
        step(q + 1, false, q + 1 <= qb, va, vb);
      }
      if (q <= whi) step(q, false, q <= qb, vb, va);
      s += dot16(acc, msk);
    }
  }
  return s;
}

// 4-cliques of one root on the matrix cores: the sum over DAG edges x -> y
// inside L of |R[x] & R[y]| (each triangle x < y < z of L counted at its
// first edge): the block-pair form of tri_mfma1 with every member of L in
// the blocks, so the members are the row positions themselves.  Waves take
// the Y blocks wv, wv + G, ...; block Y's positions lie in word Y / 2, so
// the edge mask is one word of row x per block pair.
template <class RowT>
__device__ uint64_t tri4_mfma(const RowT& R, int d_, int W_, int wv_, int G) {
  const int d = uni(d_), whi = uni(W_) - 1, wv = uni(wv_), zr = d - 1;
  const int ln = lane_id();
  const int col = ln & 31, half = ln >> 5;
  const int T = (d + 31) >> 5;
  uint64_t s = 0;
  for (int Y = wv; Y < T; Y += G) {
    const int q0 = (32 * Y + 1) >> 6;  // R[y] is empty up to the block's first member
    const int ay = 32 * Y + col;
    const int ry = ay < d ? ay : zr;
    for (int X = 0; X <= Y; ++X) {
      const int bx = 32 * X + col;
      const int rx = bx < d ? bx : zr;
      const RowCur ky = R.cur(ry, half), kx = R.cur(rx, half);
      uint32_t va[2] = {R.at(ky, q0), R.at(kx, q0)}, vb[2];
      auto step = [&](int q, const v16f32& c, const uint32_t(&v)[2], uint32_t(&nx)[2]) -> v16f32 {
        const uint32_t sa = v[0], sb = v[1];
        const int wn = min(q + 1, whi);
        nx[0] = R.at(ky, wn);
        nx[1] = R.at(kx, wn);
        return mma_fp4(enc_a(sa), enc_b(sb), c);
      };
      v16f32 acc = step(q0, zero16(), va, vb);
      int q = q0 + 1;
      for (; q + 1 <= whi; q += 2) {
        acc = step(q, acc, vb, va);
        acc = step(q + 1, acc, va, vb);
      }
      if (q <= whi) acc = step(q, acc, vb, va);
      // edge x -> y: bit y of R[x]; accumulator 4g + e holds row 8g + 4 half + e
      float f = 0.f;
      if (bx < d) {
        const uint64_t rw = R(bx, Y >> 1);
#pragma unroll
        for (int g = 0; g < 4; ++g) {
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int yy = 32 * Y + 8 * g + 4 * half + e;
            if (yy < d && ((rw >> (yy & 63)) & 1ull)) f += acc[4 * g + e];
          }
        }
      }
      s += (uint32_t)f;
    }
  }
  return s;
}

// The same count with 2 x 2 tiles of block pairs (the 2048-wide roots, whose
// rows are read from global memory, so each loaded operand feeding two
// products pays): each word's four operand slices feed four independent
// count accumulators.  The edge
// mask is the product E_Y R_X^T of one-hot rows at the Y members' own
// positions, accumulated on the words that hold them (q0 .. qm).  Pairs of
// the tile outside the triangle X <= Y, or with absent blocks, have an empty
// mask or empty operands, so they add zero.
template <class RowT>
__device__ uint64_t tri_mfma(const RowT& R, const int16_t* J, int n_, int i_, int whi_, int zr_) {
  const int n = uni(n_), i = uni(i_), whi = uni(whi_), zr = uni(zr_);
  const int ln = lane_id();
  const int col = ln & 31, half = ln >> 5;
  const int T = (n + 31) >> 5;
  auto member = [&](int B) -> int {
    const int a = 32 * B + col;
    return a < n ? (int)J[a] : -1;
  };
  uint64_t s = 0;
  for (int Y0 = 0; Y0 < T; Y0 += 2) {
    const int y0 = member(Y0), y1 = member(Y0 + 1);
    const int ry0 = y0 >= 0 ? y0 : zr, ry1 = y1 >= 0 ? y1 : zr;
    // words q0 .. qm hold the members of the two Y blocks
    const int q0 = uni(J[32 * Y0]) >> 6, qm = uni(J[min(n - 1, 32 * Y0 + 63)]) >> 6;
    const int Xend = min(Y0 + 1, T - 1);
    for (int X0 = 0; X0 <= Xend; X0 += 2) {
      const int x0 = member(X0), x1 = X0 + 1 <= Xend ? member(X0 + 1) : -1;
      const int rx0 = x0 >= 0 ? x0 : zr, rx1 = x1 >= 0 ? x1 : zr;
      v16f32 c00, c01, c10, c11, m00, m01, m10, m11;
      // word q + 1's row halves are loaded before word q's table reads
      // (two register sets alternate, so the loop moves no values)
      const RowCur ki = R.cur(i, half), ky0 = R.cur(ry0, half), ky1 = R.cur(ry1, half);
      const RowCur kx0 = R.cur(rx0, half), kx1 = R.cur(rx1, half);
      uint32_t va[5] = {R.at(ki, q0), R.at(ky0, q0), R.at(ky1, q0), R.at(kx0, q0), R.at(kx1, q0)}, vb[5];
      auto step = [&](int q, bool first, bool with_mask, const uint32_t(&v)[5], uint32_t(&nx)[5]) {
        const uint32_t sa0 = v[1] & v[0], sa1 = v[2] & v[0], sb0 = v[3], sb1 = v[4];
        const int wn = min(q + 1, whi);
        nx[0] = R.at(ki, wn);
        nx[1] = R.at(ky0, wn);
        nx[2] = R.at(ky1, wn);
        nx[3] = R.at(kx0, wn);
        nx[4] = R.at(kx1, wn);
        const v8i32 a0 = enc_a(sa0), a1 = enc_a(sa1);
        const v8i32 b0 = enc_b(sb0), b1 = enc_b(sb1);
        const v16f32 z = zero16();
        c00 = mma_fp4(a0, b0, first ? z : c00);
        c01 = mma_fp4(a0, b1, first ? z : c01);
        c10 = mma_fp4(a1, b0, first ? z : c10);
        c11 = mma_fp4(a1, b1, first ? z : c11);
        if (with_mask) {
          const v8i32 e0 = one_hot_fp4(y0, q, half), e1 = one_hot_fp4(y1, q, half);
          m00 = mma_fp4(e0, b0, first ? z : m00);
          m01 = mma_fp4(e0, b1, first ? z : m01);
          m10 = mma_fp4(e1, b0, first ? z : m10);
          m11 = mma_fp4(e1, b1, first ? z : m11);
        }
      };
      step(q0, true, true, va, vb);  // word q0 holds the first Y member
      int q = q0 + 1;
      for (; q + 1 <= whi; q += 2) {
        step(q, false, q <= qm, vb, va);
        step(q + 1, false, q + 1 <= qm, va, vb);
      }
      if (q <= whi) step(q, false, q <= qm, vb, va);
      s += (uint64_t)dot16(c00, m00) + dot16(c01, m01) + dot16(c10, m10) + dot16(c11, m11);
    }
  }
  return s;
}

// Member table of 16-bit positions into L, power-of-two size with at least
// four slots per member (load <= 1/4): Fibonacci hash, linear probing; the
// stored position is confirmed against L.
__device__ __forceinline__ int hfind16(const int16_t* h16, const int32_t* L, int mask, int shift, int32_t y) {
  int h = (int)(((uint32_t)y * 2654435761u) >> shift);
  for (;;) {
    const int t = h16[h];
    if (t < 0) return -1;
    if (L[t] == y) return t;
    h = (h + 1) & mask;
  }
}

// LDS layout of one root slot.  The build-phase arrays (segment offsets,
// list bases, long-list degrees) and the count-phase arrays (compacted row
// lists, deferred rows, the compressed bitmap) share one region.
template <int G, int WMAX, int K, bool BYTES, bool LROWS, bool CNT = true, bool BUILD = true>
struct Lay {
  static constexpr int DM = 64 * WMAX;
  static constexpr bool ROWS = K >= 4;
  static constexpr bool MULTI = WMAX > 1 && ROWS;
  static constexpr size_t al(size_t x) { return (x + 15) & ~size_t(15); }
  static constexpr size_t L_OFF = 0;
  static constexpr size_t W_OFF = al(L_OFF + (BUILD ? 4 * DM : 0));  // the root list (build only)
  static constexpr size_t R_OFF = al(W_OFF + (BYTES ? 4 * DM : 0));
  static constexpr size_t U_OFF = al(R_OFF + ((LROWS && ROWS) ? (size_t)8 * DM * WMAX : 0));
  // build view
  static constexpr size_t B_OFFS = 0;
  static constexpr size_t B_UB = al(B_OFFS + 4 * (DM + 1));
  // member lookup: binary search over L for one-word roots, else the
  // 16-bit position hash with 4 * DM slots
  static constexpr bool HASH = WMAX > 1;
  static constexpr int HCAP = HASH ? 4 * DM : 0;
  static constexpr size_t HBYTES = 2 * (size_t)HCAP;
  static constexpr size_t KBYTES = MULTI ? 4 * (size_t)DM : 0;
  static constexpr size_t B_HFL = al(B_UB + 8 * DM);                    // u32 [DM/32] long-list flags
  static constexpr size_t B_KEY = al(B_HFL + 4 * ((DM + 31) / 32));      // int32 [DM] order keys, then int16 [HCAP] hash
  static constexpr size_t B_RNK = al(B_KEY + (HBYTES > KBYTES ? HBYTES : KBYTES));  // int16 [DM] rank in DAG order
  static constexpr size_t B_END = al(B_RNK + (MULTI ? 2 * DM : 0));
  // count view, one slot of C_PW bytes per wave: the positions J of S
  // (int16 [DM]) and the successor bitmap Rc (u64 [64])
  static constexpr size_t C_RC = al(2 * (size_t)DM);
  static constexpr size_t C_PW = MULTI ? C_RC + 8 * 64 : 0;
  static constexpr size_t C_END = al((size_t)G * C_PW);
  static constexpr size_t U_SIZE = (!CNT || B_END > C_END) ? B_END : C_END;  // CNT: the count view is used
  static constexpr size_t SLOT = al(U_OFF + U_SIZE);
};

// Process one root v0 (list N+(v0) = ci[b0 .. b0 + d)) with a group of G
// waves: build the bitmap rows in `rowp` and add the root's k-cliques (or,
// with BYTES, its work-model bytes) to `acc`.  `ar` is the group's LDS
// arena of Lay<G, WMAX, K, BYTES, LROWS>::SLOT bytes.
template <int G, int WMAX, int K, bool BYTES, bool LROWS, int PHASE = 0>
__device__ void process_root(const int64_t* __restrict__ rp, const int32_t* __restrict__ ci,
                             const int32_t* __restrict__ ord, int64_t b0, int d, int ratio, unsigned char* ar,
                             uint64_t* rowp, int32_t* s_wtot, uint64_t& acc, uint32_t* grow = nullptr) {
  // PHASE 0: build and count; 1: build only (rows left in rowp); 2: count
  // only (rowp already holds the rows of this root).  grow: a global row
  // counter (zero before) that several groups counting the same root share;
  // null: the group's own counter in LDS
  static_assert(PHASE == 0 || (!BYTES && K >= 4), "split phases are for counting with rows");
  // Exactness of the fp4 matrix-core counts (tri_mfma1, tri_mfma, tri4_mfma):
  // every product of two set e2m1 elements is exactly 1, an accumulator entry
  // counts at most 64 * WMAX positions, and the epilogues add 16 entries in
  // f32.  Both stay integers below 2^24 (f32's exact range) only while
  // 16 * 64 * WMAX < 2^24; the row-width tiers (WMAX <= 32) are far inside.
  static_assert(16 * 64 * (long)WMAX < (1l << 24), "fp4 counts would leave f32's exact integer range");
  static_assert(64 * WMAX <= 2048, "rows wider than 2048 positions: the 16-bit position hash and the "
                                   "exactness notes of the fp4 epilogues assume at most 2048");
  using LY = Lay<G, WMAX, K, BYTES, LROWS, PHASE != 1, PHASE != 2>;
  constexpr int DM = LY::DM;
  constexpr bool NEED_ROWS = LY::ROWS;
  const int r = Grp<G>::rank();
  const int wv = Grp<G>::wave();
  const int ln = lane_id();
  int32_t* L = reinterpret_cast<int32_t*>(ar + LY::L_OFF);
  uint32_t* wL = reinterpret_cast<uint32_t*>(ar + LY::W_OFF);
  unsigned char* U = ar + LY::U_OFF;
  int32_t* off = reinterpret_cast<int32_t*>(U + LY::B_OFFS);
  int64_t* ub = reinterpret_cast<int64_t*>(U + LY::B_UB);
  uint32_t* hfl = reinterpret_cast<uint32_t*>(U + LY::B_HFL);
  int32_t* key = reinterpret_cast<int32_t*>(U + LY::B_KEY);
  int16_t* rnk = reinterpret_cast<int16_t*>(U + LY::B_RNK);
  int16_t* h16 = reinterpret_cast<int16_t*>(key);  // after ranking, the key region holds the hash table
  // Multi-word roots number their members in DAG order (rank of ord[]), so
  // every row only has bits above its own position.
  constexpr bool RANK = LY::MULTI;
  const int W = (d + 63) >> 6;
  const RowLayout lay(W);
  if constexpr (PHASE != 2) {
  if constexpr (BYTES) {
    if (r == 0) acc += 16 + 4 * (uint64_t)d;
  }
  // ---- stage root list and segment descriptors
  for (int t = r; t < d; t += Grp<G>::size) L[t] = ci[b0 + t];
  for (int t = r; t < (d + 31) / 32; t += Grp<G>::size) hfl[t] = 0u;
  Grp<G>::sync();
  for (int t = r; t < d; t += Grp<G>::size) {
    int32_t u = L[t];
    int64_t s = rp[u];
    int du = (int)(rp[u + 1] - s);
    bool heavy = du > 64 && du > ratio * d;
    ub[t] = s;
    off[t] = heavy ? 0 : du;
    if (heavy) atomicOr(&hfl[t >> 5], 1u << (t & 31));
    if constexpr (RANK) key[t] = ord[u];
    if constexpr (BYTES) acc += 16 + 4 * (uint64_t)du;  // f(v1) of the edge (v0, L[t])
  }
  if constexpr (RANK) {
    Grp<G>::sync();
    for (int t = r; t < d; t += Grp<G>::size) {
      const int32_t kt = key[t];
      int c = 0;
      for (int s2 = 0; s2 < d; ++s2) c += key[s2] < kt;
      rnk[t] = (int16_t)c;
    }
  }
  Grp<G>::sync();
  if constexpr (BYTES) {
    for (int t = r; t < d; t += Grp<G>::size) {
      const int pos = RANK ? (int)rnk[t] : t;
      wL[pos] = 16u + 4u * (uint32_t)(rp[L[t] + 1] - ub[t]);
    }
  }
  // hash table of the members: slot -> position in L (power of two >= 4d
  // slots, so probes stay short and an empty slot always exists)
  int hs = 0, hshift = 0;
  if constexpr (LY::HASH) {
    int lg = 6;
    while ((1 << lg) < 4 * d) ++lg;
    hs = 1 << lg;
    hshift = 32 - lg;
    uint32_t* hw = reinterpret_cast<uint32_t*>(h16);
    for (int t = r; t < (hs >> 1); t += Grp<G>::size) hw[t] = 0xffffffffu;
    Grp<G>::sync();
    for (int t = r; t < d; t += Grp<G>::size) {
      // two 16-bit slots per word: claim the first empty one by CAS on the word
      int h = (int)(((uint32_t)L[t] * 2654435761u) >> hshift);
      volatile uint32_t* vw = hw;
      uint32_t old = vw[h >> 1];
      for (;;) {
        const int sh = (h & 1) * 16;
        if (((old >> sh) & 0xffffu) != 0xffffu) {
          h = (h + 1) & (hs - 1);
          old = vw[h >> 1];
          continue;
        }
        const uint32_t nv = (old & ~(0xffffu << sh)) | ((uint32_t)t << sh);
        const uint32_t prev = atomicCAS(&hw[h >> 1], old, nv);
        if (prev == old) break;
        old = prev;
      }
    }
  }
  if constexpr (NEED_ROWS) {
    for (int t = r; t < d * lay.S; t += Grp<G>::size) rowp[t] = 0;
  }
  if constexpr (!LROWS && NEED_ROWS) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  Grp<G>::sync();
  if constexpr (K == 3 && BYTES) return;  // TC work model needs no build
  group_excl_scan<G>(off, d, s_wtot);
  const int T = off[d];
  // light lists: element f of the stream sits at ci[ub[i] + f] (one LDS read
  // per element for its address); long lists keep their own base
  for (int t = r; t < d; t += Grp<G>::size)
    if (!((hfl[t >> 5] >> (t & 31)) & 1u)) ub[t] -= off[t];
  Grp<G>::sync();
  // ---- flattened stream of the light lists: each wave takes one contiguous
  // share of the elements and its lanes step through it 64 at a time
  // (consecutive lanes read consecutive addresses).  A lane's segment index
  // only moves forward, by the few lists 64 elements span, and eight loads
  // are in flight at once (twelve where the element feeds a row: measured
  // faster there, slower for the triangle count's hit counter).
  {
    constexpr int UF = K == 3 ? 8 : 12;
    const int per = ((T + G - 1) / G + 63) & ~63;
    const int fb = wv * per, fe = min(T, fb + per);
    int i = fb + ln < fe ? seg_search(off, d, fb + ln) : 0;
    for (int f0 = fb + ln; f0 < fe; f0 += UF * 64) {
      int32_t yv[UF];
      int iv[UF];
#pragma unroll
      for (int u = 0; u < UF; ++u) {
        const int f = f0 + u * 64;
        iv[u] = -1;
        if (f < fe) {
          i = seg_advance(off, d, i, f);
          iv[u] = i;
          yv[u] = ci[ub[i] + f];
        }
      }
#pragma unroll
      for (int u = 0; u < UF; ++u) {
        if (iv[u] < 0) continue;
        const int32_t y = yv[u];
        int rpb;  // bit position of y in the rows, -1 if y is not a member
        if constexpr (LY::HASH) {
          rpb = hfind16(h16, L, hs - 1, hshift, y);
          if (RANK && rpb >= 0) rpb = rnk[rpb];
        } else {
          const int p = lower_bound_dev(L, d, y);
          rpb = (p < d && L[p] == y) ? p : -1;
        }
        if (rpb >= 0) {
          if constexpr (K == 3) {
            ++acc;
          } else {
            const int ri = RANK ? (int)rnk[iv[u]] : iv[u];
            atomicOr((unsigned long long*)&rowp[lay(ri, rpb >> 6)], 1ull << (rpb & 63));
          }
        }
      }
    }
  }
  // ---- long lists: binary search every element of L in them.  The
  // (list, 64-member chunk) pairs are dealt to the waves round-robin, so a
  // root with fewer members than 64 G still keeps every wave busy.
  const int nch = (d + 63) >> 6;
  int kk = 0;  // pairs dealt so far (the same on every wave)
  for (int i = 0; i < d; ++i) {
    if (!((hfl[i >> 5] >> (i & 31)) & 1u)) continue;
    const int du = (int)(rp[L[i] + 1] - ub[i]);
    const int32_t* lst = ci + ub[i];
    const int ri = RANK ? (int)rnk[i] : i;
    const int c0 = ((wv - kk) % G + G) % G;  // first chunk c with (kk + c) % G == wv
    kk += nch;
    for (int t0 = c0 * 64; t0 < d; t0 += Grp<G>::size) {
      int t = t0 + ln;
      bool f = t < d && bsearch_global(lst, du, L[t]);
      if constexpr (K == 3) {
        uint64_t m = __ballot(f);
        if (ln == 0) acc += (uint64_t)__popcll(m);
      } else if constexpr (RANK) {
        if (f) {
          const int rt = rnk[t];
          atomicOr((unsigned long long*)&rowp[lay(ri, rt >> 6)], 1ull << (rt & 63));
        }
      } else {
        uint64_t m = __ballot(f);
        if (ln == 0 && m) atomicOr((unsigned long long*)&rowp[lay(ri, t0 >> 6)], m);
      }
    }
  }
  if constexpr (!LROWS && NEED_ROWS) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  if (G > 1 && r == 0) s_wtot[G] = 0;  // row counter of the count phase (multi-word rows)
  Grp<G>::sync();
  if constexpr (!LROWS && NEED_ROWS) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }  // build
  if constexpr (K == 3 || PHASE == 1) return;
  // ---- count (the build-phase region U is now free)
  if constexpr (!LY::MULTI) {
    // one word per row, lane i owns row i
    if (r < d) {
      const uint64_t Ri = rowp[r];
      if constexpr (BYTES) acc += rec1<K - 3, true>(Ri, rowp, wL);
      else acc += rec1<K - 2, false>(Ri, rowp, wL);
    }
  } else {
    using RowT = typename std::conditional<LROWS, RowsLds, RowsGlobal>::type;
    const RowT R{rowp, lay};
    // 4-cliques of the 512-wide and wider roots as one masked block product
    // over the whole root (P13 on R-MAT 22, one quarter of the roots: split
    // count 3.2 -> 0.75 ms; 512-wide tier 10.8 -> 9.7 ms with its registers
    // held to the 6 waves per SIMD its LDS allows, 13.0 ms without); the
    // 256-wide and narrower tiers keep the bitwise count
    if constexpr (K == 4 && !BYTES && WMAX >= 8) {
      acc += tri4_mfma(R, d, W, wv, G);
    } else {
    unsigned char* slot = U + (size_t)wv * LY::C_PW;
    int16_t* J = reinterpret_cast<int16_t*>(slot);
    uint64_t* Rc = reinterpret_cast<uint64_t*>(slot + LY::C_RC);
    // waves take rows in order from a shared counter (s_wtot[G], zeroed
    // before the group's last barrier): early rows in DAG order hold the most
    // work, so a fixed interleave leaves waves idle at the end.  4-cliques
    // on the narrower tiers have too little work per row to pay for it.
    constexpr bool DYN = G > 1 && (K >= 5 || WMAX >= 16);
    // 5-cliques of every row (also |S| <= 64) on the matrix cores from 256
    // wide up: the successor bitmap below leaves lanes idle on rows of few
    // members (measured 79.8 -> 75.3 ms for one P14 query on R-MAT scale
    // 22); the 128-wide tier is faster without (6.4 vs 8.2 ms)
    constexpr bool MMA5 = K == 5 && !BYTES && WMAX >= 4;
    auto next_row = [&](int cur) -> int {
      if constexpr (!DYN) {
        return cur + G;
      } else {
        int t = 0;
        if (grow) {
          if (ln == 0) t = (int)atomicAdd(grow, 1u);
        } else {
          // the counter's LDS address from a zero made here (the empty asm
          // keeps it from being hoisted): a loop-invariant address register
          // would be spilled across the row's matrix-core work and reloaded
          int z = 0;
          asm volatile("" : "+v"(z));
          if (ln == 0) t = atomicAdd(s_wtot + G + z, 1);
        }
        return __builtin_amdgcn_readfirstlane(t);
      }
    };
    // the lane's share of this root's 5-cliques on the 256- and 512-wide
    // tiers, in 32 bits: at most 512 rows of at most 136 block pairs, each
    // adding at most 16 entries of at most 512, below 2^30.  It is kept in
    // the wave's successor-bitmap slot (unused by the matrix-core count) and
    // added to without a return value: a sum held in registers across the
    // row loop costs those tiers a register spill and reload per row.
    uint32_t* racc = reinterpret_cast<uint32_t*>(Rc) + ln;
    if constexpr (MMA5 && WMAX <= 8) *racc = 0u;
    static_assert(!MMA5 || WMAX > 8 || (long)(64 * WMAX) * (WMAX * (2 * WMAX + 1)) * 16 * (64 * WMAX) < (1l << 32),
                  "the 32-bit root share of the narrow 5-clique tiers could overflow");
    // the next row is taken before this row's work, so its counter round
    // trip overlaps the work
    int inext = 0;
    for (int i = DYN ? next_row(0) : wv; i < d; i = inext) {
      inext = next_row(i);
      // population and nonzero word window of row i (W <= 32 words)
      const uint64_t myword = ln < W ? R(i, ln) : 0ull;
      const uint64_t nz = __ballot(myword != 0ull);
      if (!nz) continue;
      const int wlo = __builtin_ctzll(nz);
      const int whi = 63 - __builtin_clzll(nz);
      // compact S = row i into J (ascending positions)
      int n = 0;
      for (int w = wlo; w <= whi; ++w) {
        const uint64_t word = R(i, w);
        if ((word >> ln) & 1ull) {
          // set bits of the word below this lane's bit
          const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(word >> 32),
                                                           __builtin_amdgcn_mbcnt_lo((uint32_t)word, 0u));
          J[n + below] = (int16_t)((w << 6) | ln);
        }
        n += __popcll(word);
      }
      __syncwarp();
      if constexpr (K == 4) {
        for (int t = ln; t < n; t += 64) {
          const int j = J[t];
          if constexpr (BYTES) {
            acc += wL[j];
          } else {
            uint64_t c = 0;
            for (int w = max(wlo, (j + 1) >> 6); w <= whi; ++w) c += __popcll(R(i, w) & R(j, w));
            acc += c;
          }
        }
      } else if (!MMA5 && n <= 64) {
        // compressed |S| x |S| successor bitmap, one word per member
        uint64_t myrow = 0;
        const int js = ln < n ? J[ln] : 0;
        int curw = -1;
        uint64_t word = 0;
        for (int t = 0; t < n; ++t) {
          const int x = J[t];
          if (ln < n) {
            if ((x >> 6) != curw) {
              curw = x >> 6;
              word = R(js, curw);
            }
            myrow |= ((word >> (x & 63)) & 1ull) << t;
          }
        }
        if (ln < n) Rc[ln] = myrow;
        __syncwarp();
        if (ln < n) {
          // weights of compressed members go through J back to L positions
          if constexpr (BYTES) acc += wL[js] + rec1<K - 4, true, int16_t>(myrow, Rc, wL, J);
          else acc += rec1<K - 3, false>(myrow, Rc, wL);
        }
      } else {
        if constexpr (MMA5) {
          // tiles of block pairs pay for their registers only on the 2048-wide
          // roots (rows in global memory); the edge mask is a product from
          // 512 wide up
          if constexpr (WMAX >= 32) acc += tri_mfma(R, J, n, i, whi, d - 1);
          else if constexpr (WMAX >= 16) acc += tri_mfma1(R, J, n, i, whi, d - 1);
          else if constexpr (WMAX >= 8)
            __hip_atomic_fetch_add(racc, (uint32_t)tri_mfma1(R, J, n, i, whi, d - 1), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WAVEFRONT);
          else
            __hip_atomic_fetch_add(racc, (uint32_t)tri_mfma1_bits(R, J, n, i, whi, d - 1), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WAVEFRONT);
        } else if constexpr (K == 5 && WMAX <= 16) {
          // 5-cliques: triangles inside S.  One flat loop per lane over a
          // work list shared through an LDS counter: a step either takes the
          // next member j (candidates c = R[i] & R[j]) or evaluates one leaf
          // x of c (|c & R[x]|); both are the same masked word loop, so lanes
          // stay busy until the row's work is gone.
          const int nw = whi - wlo + 1;
          int* nxt = reinterpret_cast<int*>(Rc);
          if (ln == 0) *nxt = 0;
          __syncwarp();
          uint64_t ri[WMAX], c[WMAX];
#pragma unroll
          for (int w = 0; w < WMAX; ++w) {
            ri[w] = (w < nw) ? R(i, wlo + w) : 0ull;
            c[w] = 0ull;
          }
          int t = atomicAdd(nxt, 1);
          int cur = -1;       // current leaf bit in c, -1: take a new member
          int y = t < n ? J[t] : 0;
          bool grab = true;
          while (__ballot(t < n) != 0ull) {
            if (t < n) {
              const int ws = ((y + 1) >> 6) - wlo;  // row y has no bits before its own position
              uint64_t sum = 0;
              uint64_t val[WMAX];
#pragma unroll
              for (int w = 0; w < WMAX; ++w) {
                const uint64_t base = grab ? ri[w] : c[w];
                val[w] = (w >= ws && w < nw && base) ? (base & R(y, wlo + w)) : 0ull;
                sum += (uint64_t)__popcll(val[w]);
              }
              if (grab) {
#pragma unroll
                for (int w = 0; w < WMAX; ++w) c[w] = val[w];
                if constexpr (BYTES) acc += wL[y];
                cur = next_bit<WMAX>(c, nw, 0);
              } else {
                if constexpr (BYTES) acc += wL[y];
                else acc += sum;
                cur = next_bit<WMAX>(c, nw, cur + 1);
              }
              if (cur < 0) {
                t = atomicAdd(nxt, 1);
                grab = true;
                y = t < n ? J[t] : 0;
              } else {
                grab = false;
                y = wlo * 64 + cur;
              }
            }
          }
        } else
        // lanes over the members j of S; candidate set R[i] & R[j] held
        // in a register window of the row's nonzero words
        for (int t = ln; t < n; t += 64) {
          const int j = J[t];
          if constexpr (WMAX <= 16) {
            const int nw = whi - wlo + 1;
            const int ws = ((j + 1) >> 6) - wlo;  // row j has no bits before its own position
            uint64_t c[WMAX];
#pragma unroll
            for (int w = 0; w < WMAX; ++w) c[w] = (w >= ws && w < nw) ? (R(i, wlo + w) & R(j, wlo + w)) : 0ull;
            if constexpr (BYTES) acc += wL[j] + recR<K - 4, WMAX, true>(c, wlo, nw, R, wL);
            else acc += recR<K - 3, WMAX, false>(c, wlo, nw, R, wL);
          } else {
            int p2[2] = {i, j};
            if constexpr (BYTES) {
              acc += wL[j];
              acc += recW<K - 4, 2, true>(RowsGlobal{rowp, lay}, wlo, whi, p2, wL);
            } else {
              acc += recW<K - 3, 2, false>(RowsGlobal{rowp, lay}, wlo, whi, p2, wL);
            }
          }
        }
      }
      __syncwarp();
    }
    if constexpr (MMA5 && WMAX <= 8) acc += __hip_atomic_load(racc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
  }
  Grp<G>::sync();
}

}  // namespace jk
