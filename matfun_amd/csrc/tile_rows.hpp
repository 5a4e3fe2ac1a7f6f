// tile_rows.hpp -- where the up to 64 matrices of one tile (one wavefront) live in memory.
//   RunRows   consecutive matrices base .. base + nm - 1 (the C ABI's homogeneous calls): the input / F, J and H rows
//             of tile matrix j are all base + j, so every region a tile writes is one contiguous block
//   ListRows  rows listed per matrix (heterogeneous batches): input / F row m[j], J row jr[j], H row hr[j], copied
//             into LDS at the start of the tile
// The tile bodies (smf_tile.hpp, gmf_tile.hpp) take either; with RunRows they compile to the contiguous addressing,
// with ListRows each run of consecutive rows (a batch's groups) is written as one block.
#pragma once
#include "mf_device.hpp"

namespace mf {

constexpr int WAVE = 64;

struct RunRows {
  static constexpr bool RUN = true, MASKS = false, SCATTER = false;
  int64_t base;
  MF_DEV int64_t m(int j) const { return base + j; }
  MF_DEV int64_t jr(int j) const { return base + j; }
  MF_DEV int64_t hr(int j) const { return base + j; }
};

struct ListRows {
  static constexpr bool RUN = false, MASKS = false, SCATTER = false;
  const int64_t* mr;              // LDS
  const int64_t* jrr;
  const int64_t* hrr;
  MF_DEV int64_t m(int j) const { return mr[j]; }
  MF_DEV int64_t jr(int j) const { return jrr[j]; }
  MF_DEV int64_t hr(int j) const { return hrr[j]; }
};

// rows listed per matrix with less LDS (for a tile body whose own LDS leaves no room for three 64-bit rows per
// matrix): input / F rows 32-bit in LDS, J and H rows in registers (lane j holds tile matrix j's, read across lanes
// by ds_bpermute: every call of jr / hr must be made by the whole wavefront), and per region (Region below) a mask
// of where the rows continue (bit j: row j is row j - 1 plus one).  base: the tile matrix of index 0 (shift_rows).
struct ListRows32 {
  static constexpr bool RUN = false, MASKS = true, SCATTER = false;
  const int32_t* mr;              // LDS, from tile matrix base
  int64_t jl, hl;                 // this lane's tile matrix's J / H row
  int base;
  uint64_t cont[3];
  MF_DEV int64_t m(int j) const { return (int64_t)mr[j]; }
  MF_DEV int64_t jr(int j) const { return __shfl(jl, (base + j) & (WAVE - 1)); }
  MF_DEV int64_t hr(int j) const { return __shfl(hl, (base + j) & (WAVE - 1)); }
};

// rows given by a permutation (the length-sorted general-matrix path): tile matrix j is row p[j] of the input and
// of F, J and H alike; the rows in LDS (32-bit), the masks of where they continue (the same for every region) by
// ballot when they are staged
struct PermRows {
  static constexpr bool RUN = false, MASKS = true, SCATTER = true;
  const int32_t* p;               // LDS
  uint64_t cont[3];
  MF_DEV int64_t m(int j) const { return (int64_t)p[j]; }
  MF_DEV int64_t jr(int j) const { return (int64_t)p[j]; }
  MF_DEV int64_t hr(int j) const { return (int64_t)p[j]; }
};

// the barrier of one tile's wavefront: the ordering of __syncthreads (LDS writes before it visible to the wave's
// reads after it) without the workgroup's s_barrier, so that the tiles of a workgroup of several wavefronts run their
// phases independently (for a workgroup of one wavefront __syncthreads compiles to the same)
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// a copy of v / p that the compiler cannot see through (an empty asm statement): what a tile computes from it is not
// hoisted out of a loop over tiles, so a persistent kernel keeps no per-tile state live from one tile to the next
MF_DEV int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}
template <class T>
MF_DEV const T* opaque(const T* p) {
  asm volatile("" : "+s"(p));
  return p;
}

// the rows of tile matrices p0, p0 + 1, ... (a pass over part of the tile)
MF_DEV RunRows shift_rows(const RunRows& r, int p0) { return RunRows{r.base + p0}; }
MF_DEV ListRows32 shift_rows(const ListRows32& r, int p0) {
  return ListRows32{r.mr + p0, r.jl, r.hl, r.base + p0, {r.cont[0] >> p0, r.cont[1] >> p0, r.cont[2] >> p0}};
}
MF_DEV ListRows shift_rows(const ListRows& r, int p0) { return ListRows{r.mr + p0, r.jrr + p0, r.hrr + p0}; }
MF_DEV PermRows shift_rows(const PermRows& r, int p0) {
  return PermRows{r.p + p0, {r.cont[0] >> p0, r.cont[1] >> p0, r.cont[2] >> p0}};
}

// which row of a tile matrix an output region uses
enum Region { R_IN = 0, R_J = 1, R_H = 2 };
template <int REG, class R>
MF_DEV int64_t row_of(const R& r, int j) {
  return REG == R_IN ? r.m(j) : (REG == R_J ? r.jr(j) : r.hr(j));
}

// a store of the kernels' outputs (a scalar or a 16-byte vector).  MF_NT_STORE 1: non-temporal stores (the outputs
// are written once and not read back by the kernel)
#ifndef MF_NT_STORE
#define MF_NT_STORE 0
#endif
template <class T>
MF_DEV void out_store(T* p, const T& v) {
#if MF_NT_STORE
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}

// out[row(j) N + e] = get(j, e) for the count = N nm elements (t = j N + e) of the tile's matrices j = 0 .. nm-1 in
// region REG, by the wave, for consecutive rows (RunRows: one contiguous block, out + base N + t; rows listed per
// matrix take for_row_runs with copy_run / copy_run_lut).  With VEC > 1, VEC consecutive elements per lane and one
// vector store (VEC x sizeof(OT) bytes: 16 B for FP64 pairs and FP32 quadruples; the block must be aligned to it),
// the last count % VEC elements one at a time; else one element per lane.
template <int N, int VEC, int REG, class R, class OT, class Get>
MF_DEV void copy_out(OT* __restrict__ out, const R& rows, int count, int lane, Get get) {
  static_assert(R::RUN, "copy_out writes consecutive rows");
  OT* blk = out + row_of<REG>(rows, 0) * N;
  if constexpr (VEC > 1) {
    typedef OT vv __attribute__((ext_vector_type(VEC)));
    int t = VEC * lane, j = t / N, e = t - j * N;
    for (; t + VEC - 1 < count; t += VEC * WAVE) {
      vv val;
      int jq = j, eq = e;
#pragma unroll
      for (int q = 0; q < VEC; q++) {
        val[q] = get(jq, eq);
        const bool wrap = eq + 1 == N;
        eq = wrap ? 0 : eq + 1;
        jq = wrap ? jq + 1 : jq;
      }
      out_store(reinterpret_cast<vv*>(blk + t), val);
      e += VEC * WAVE;
      while (e >= N) { e -= N; j++; }
    }
    for (; t < count; t++) {                     // (the tail: at most one lane, fewer than VEC elements)
      out_store(blk + t, get(j, e));
      if (++e == N) { e = 0; j++; }
    }
  } else {
    int j = lane / N, e = lane - j * N;
    for (int t = lane; t < count; t += WAVE) {
      out_store(blk + t, get(j, e));
      e += WAVE;
      while (e >= N) { e -= N; j++; }
    }
  }
}

// the maximal runs of consecutive rows (region REG) among the tile matrices [0, cnt): fn(j0, j1) for each run
// [j0, j1), in order, by the whole wave (the runs come from a ballot: uniform).  Rows listed per matrix are mostly
// consecutive (a batch's groups of matrices), and a run's outputs form one contiguous block.
template <int REG, class R, class Fn>
__device__ void for_row_runs(const R& rows, int cnt, int lane, Fn fn) {
  uint64_t cont;
  if constexpr (R::MASKS) {
    cont = rows.cont[REG];
  } else {
    const bool c = lane > 0 && lane < cnt && row_of<REG>(rows, lane) == row_of<REG>(rows, lane - 1) + 1;
    cont = __ballot(c);
  }
  for (int j0 = 0; j0 < cnt;) {
    const uint64_t stop = j0 + 1 < 64 ? (~cont & (~0ull << (j0 + 1))) : 0ull;
    int j1 = stop ? __ffsll((long long)stop) - 1 : 64;
    j1 = j1 < cnt ? j1 : cnt;
    fn(j0, j1);
    j0 = j1;
  }
}

// o[t] = get(t / N, t % N) for t < count (one run's block), by the wave, in stores of 16 bytes aligned to 16 bytes:
// the head (the elements before the first aligned vector) and the tail one element per lane; o needs only the
// alignment of OT
template <int N, class OT, class Get>
__device__ void copy_run(OT* __restrict__ o, int count, int lane, Get get) {
  constexpr int V = 16 / (int)sizeof(OT);
  typedef OT vv __attribute__((ext_vector_type(V)));
  const int mis = (int)(((uintptr_t)o / sizeof(OT)) % V);
  const int head = (V - mis) % V < count ? (V - mis) % V : count;
  if (lane < head) out_store(o + lane, get(lane / N, lane % N));
  int t = head + V * lane, j = t / N, e = t - j * N;
  for (; t + V - 1 < count; t += V * WAVE) {
    vv val;
    int jq = j, eq = e;
#pragma unroll
    for (int q = 0; q < V; q++) {
      val[q] = get(jq, eq);
      const bool wrap = eq + 1 == N;
      eq = wrap ? 0 : eq + 1;
      jq = wrap ? jq + 1 : jq;
    }
    out_store(reinterpret_cast<vv*>(o + t), val);
    e += V * WAVE;
    while (e >= N) { e -= N; j++; }
  }
  const int t0 = head + (count - head) / V * V;
  if (lane < count - t0) out_store(o + t0 + lane, get((t0 + lane) / N, (t0 + lane) % N));
}

// Output layouts through a lookup table: element e of a matrix's output row is a value staged in LDS (the matrix's
// block: S elements per matrix) times a power of two.  Entry lut[e] (a short): bits 0..11 the byte offset of that
// value from the matrix's block, bits 12..15 the exponent x (signed: -8 .. 7); out = value 2^x.  Entries N .. N +
// LUT_PAD - 1 describe elements 0 .. LUT_PAD - 1 of the next matrix (offset + S elements), for vector stores that
// run across a matrix boundary.
constexpr int LUT_PAD = 3;                                   // (vectors of up to 4 elements)
template <class OT>
MF_DEV short lut_entry(int src, int x) {                     // src: staged element (< 4096 / sizeof(OT))
  return (short)(unsigned short)((src * (int)sizeof(OT)) | ((x & 15) << 12));
}
template <class OT>
MF_DEV OT lut_value(const unsigned char* blk, int u) {       // (ldexp: the exact product by 2^x)
  const OT v = *reinterpret_cast<const OT*>(blk + (u & 0xfff));
  if constexpr (sizeof(OT) == 4) return __builtin_ldexpf(v, u >> 12);
  else return __builtin_ldexp(v, u >> 12);
}

// one run's block through the lookup table: o[t] = lut_value(block of run matrix t / N, lut[t % N]), blocks at st8 +
// (t / N) S elements; 16-byte stores after an aligned head (as copy_run), the table's padding for vectors across a
// matrix boundary (as copy_out_lut)
template <int N, int S, class OT>
__device__ void copy_run_lut(OT* __restrict__ o, int count, int lane, const unsigned char* st8, const short* lut) {
  constexpr int V = 16 / (int)sizeof(OT);
  static_assert(V - 1 <= LUT_PAD && V <= N, "vector longer than the table's padding or a row");
  typedef OT vv __attribute__((ext_vector_type(V)));
  const int mis = (int)(((uintptr_t)o / sizeof(OT)) % V);
  const int head = (V - mis) % V < count ? (V - mis) % V : count;
  if (lane < head) out_store(o + lane, lut_value<OT>(st8, lut[lane]));          // (head < V <= N: the first matrix)
  int t = head + V * lane, j = t / N, e = t - j * N;
  const unsigned char* b = st8 + j * S * (int)sizeof(OT);
  for (; t + V - 1 < count; t += V * WAVE) {
    vv val;
#pragma unroll
    for (int q = 0; q < V; q++) val[q] = lut_value<OT>(b, lut[e + q]);
    out_store(reinterpret_cast<vv*>(o + t), val);
    e += V * WAVE;
    while (e >= N) { e -= N; b += S * (int)sizeof(OT); }
  }
  const int t0 = head + (count - head) / V * V;
  if (lane < count - t0) {
    const int tt = t0 + lane, jj = tt / N;
    out_store(o + tt, lut_value<OT>(st8 + jj * S * (int)sizeof(OT), lut[tt - jj * N]));
  }
}

// copy_out of a layout with a lookup table: out[row(j) N + e] = lut_value(staged block of matrix j, lut[e]), blocks
// at st + j S.  FP32 outputs, RunRows, VEC > 1: each lane keeps the offset e of its vector's first element and a
// pointer to that element's block; the table's entries past N cover the elements that fall into the next matrix, so
// the per-element work is two LDS reads, the address sum and the scaling (the mixed kernels are limited by this
// VALU work).  FP64 outputs take copy_out's per-element form: with the FP64 kernels' six wavefronts per CU and the
// store rate their bound, it measured 2-3% faster on MI355X than the short loop (same results either way).
template <int N, int S, int VEC, int REG, class R, class OT>
MF_DEV void copy_out_lut(OT* __restrict__ out, const R& rows, int count, int lane, const OT* st, const short* lut) {
  static_assert(VEC - 1 <= LUT_PAD, "vector longer than the table's padding");
  const unsigned char* st8 = reinterpret_cast<const unsigned char*>(st);
  if constexpr (R::RUN && VEC > 1 && sizeof(OT) == 4) {
    typedef OT vv __attribute__((ext_vector_type(VEC)));
    OT* blk = out + row_of<REG>(rows, 0) * N;
    int t = VEC * lane, j = t / N, e = t - j * N;
    const unsigned char* b = st8 + j * S * (int)sizeof(OT);
    for (; t + VEC - 1 < count; t += VEC * WAVE) {
      vv val;
#pragma unroll
      for (int q = 0; q < VEC; q++) val[q] = lut_value<OT>(b, lut[e + q]);
      out_store(reinterpret_cast<vv*>(blk + t), val);
      e += VEC * WAVE;
      if constexpr (N >= VEC * WAVE) {           // (at most one matrix boundary per step: selects, no division)
        const bool w = e >= N;
        e = w ? e - N : e;
        b = w ? b + S * (int)sizeof(OT) : b;
      } else {
        while (e >= N) { e -= N; b += S * (int)sizeof(OT); }
      }
    }
    for (; t < count; t++) {                     // (the tail: at most one lane, fewer than VEC elements)
      out_store(blk + t, lut_value<OT>(b, lut[e]));
      if (++e == N) { e = 0; b += S * (int)sizeof(OT); }
    }
  } else if constexpr (!R::RUN) {
    // rows listed: each run of consecutive rows as one block (a run of one matrix: its row alone)
    for_row_runs<REG>(rows, count / N, lane, [&](int j0, int j1) {
      copy_run_lut<N, S, OT>(out + row_of<REG>(rows, j0) * N, (j1 - j0) * N, lane, st8 + j0 * S * (int)sizeof(OT),
                             lut);
    });
  } else {
    copy_out<N, VEC, REG>(out, rows, count, lane,
                          [&](int j, int e) { return lut_value<OT>(st8 + j * S * (int)sizeof(OT), lut[e]); });
  }
}

// out[0 .. count) = src[0 .. count) (src in LDS), by the wave, in stores of VEC elements aligned to VEC sizeof(OT)
// bytes: the head (the elements before the first aligned vector) and the tail one element per lane, the rest in
// vectors; out needs only the alignment of OT
template <int VEC, class OT>
MF_DEV void copy_block(OT* __restrict__ out, const OT* src, int count, int lane) {
  const int mis = (int)(((uintptr_t)out / sizeof(OT)) % VEC);
  const int head = (VEC - mis) % VEC < count ? (VEC - mis) % VEC : count;
  if (lane < head) out_store(out + lane, src[lane]);
  const int nv = (count - head) / VEC;
  if constexpr (VEC > 1) {
    typedef OT vv __attribute__((ext_vector_type(VEC)));
    for (int v = lane; v < nv; v += WAVE) {
      vv val;
#pragma unroll
      for (int q = 0; q < VEC; q++) val[q] = src[head + VEC * v + q];
      out_store(reinterpret_cast<vv*>(out + head + VEC * v), val);
    }
  } else {
    for (int v = lane; v < nv; v += WAVE) out_store(out + head + v, src[head + v]);
  }
  const int t0 = head + VEC * nv;
  if (lane < count - t0) out_store(out + t0 + lane, src[t0 + lane]);
}

// out[row(j) N + e] = src[j N + e] for the count = N nm staged elements of tile matrices whose rows are scattered
// (permuted rows: runs of one matrix): one element per lane and store, all lanes busy, consecutive lanes on
// consecutive elements of a row -- instead of a store loop per run (measured on MI355X, length-sorted general
// matrices: order 1 15%, mixed order 2 4% faster than the loop over runs)
template <int N, int REG, class R, class OT>
MF_DEV void copy_rows_scattered(OT* __restrict__ out, const R& rows, const OT* src, int count, int lane) {
  int j = lane / N, e = lane - j * N;
  for (int t = lane; t < count; t += WAVE) {
    out_store(out + row_of<REG>(rows, j) * N + e, src[t]);
    e += WAVE;
    while (e >= N) { e -= N; j++; }
  }
}

// the tile's inputs (NIN doubles per matrix) into LDS, coalesced
template <int NIN, class R>
MF_DEV void load_in(double* __restrict__ xin, const double* __restrict__ x, const R& rows, int nm, int lane) {
  if (R::RUN) {
    const double* blk = x + rows.m(0) * NIN;
    for (int t = lane; t < NIN * nm; t += WAVE) xin[t] = blk[t];
  } else {
    int j = lane / NIN, e = lane - j * NIN;
    for (int t = lane; t < NIN * nm; t += WAVE) {
      xin[t] = x[rows.m(j) * NIN + e];
      e += WAVE;
      while (e >= NIN) { e -= NIN; j++; }
    }
  }
}

// the reverse of load_in: the tile's inputs from LDS back to x, coalesced
template <int NIN, class R>
MF_DEV void store_in(double* __restrict__ x, const double* __restrict__ xin, const R& rows, int nm, int lane) {
  if (R::RUN) {
    double* blk = x + rows.m(0) * NIN;
    for (int t = lane; t < NIN * nm; t += WAVE) blk[t] = xin[t];
  } else {
    int j = lane / NIN, e = lane - j * NIN;
    for (int t = lane; t < NIN * nm; t += WAVE) {
      x[rows.m(j) * NIN + e] = xin[t];
      e += WAVE;
      while (e >= NIN) { e -= NIN; j++; }
    }
  }
}

// A transform of each tile matrix's inputs before it is evaluated, the transformed inputs also written back over
// the input rows (the heterogeneous batch's Newton perturbation).  Pre: active() (uniform), apply(a, row) on one
// lane's 9 inputs, and x, where the inputs are written back.  NoPre: none (the evaluation entry points).
struct NoPre {
  double* x = nullptr;
  MF_DEV bool active() const { return false; }
  MF_DEV void apply(double*, int64_t) const {}
};

// pre applied to the tile's matrices (lane j: a = its 9 inputs, already read from xin, which held the tile's
// inputs; its row rows.m(j)), the results written back to pre.x through xin; the wave's barriers included
template <class Pre, class R>
MF_DEV void pre_inputs(const Pre& pre, double* a, double* xin, const R& rows, int nm, int lane) {
  if (!pre.active()) return;
  wave_sync();                                     // (every lane has read its inputs from xin)
  if (lane < nm) {
    pre.apply(a, rows.m(lane));
#pragma unroll
    for (int k = 0; k < 9; k++) xin[9 * lane + k] = a[k];
  }
  wave_sync();
  store_in<9>(pre.x, xin, rows, nm, lane);
  wave_sync();                                     // (xin free again)
}

}  // namespace mf
