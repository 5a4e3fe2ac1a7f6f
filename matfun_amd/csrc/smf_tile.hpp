// smf_tile.hpp -- one tile (up to 64 symmetric 3x3 matrices, one wavefront) of the symmetric kernels:
//   phase A, one matrix per lane: coalesced input through LDS, eigen-decomposition, F, the 21 distinct J values,
//            the 10 second divided differences; F / J staged in LDS and written as one contiguous region per wave
//   phase B, six lanes per matrix (one per Voigt row of H), HPP matrices per pass (the LDS store tile: 10, 8 or 5):
//            H rows from V and D2 staged in LDS, gathered into the API layout through a lookup table and written
//            contiguously (a pass covers HPP x 5.8 KB of consecutive output).  SEP: V / D2 in their own LDS block,
//            staged at the end of phase A; else staged after the F / J stores into the phase-A region (less LDS per
//            wavefront, one more barrier, V / D2 held in registers across the stores)
// Every store instruction of the wave writes 64 consecutive elements: the output (96 B .. 5.8 KB per matrix) bounds
// the kernel by HBM write bandwidth, which needs whole cache lines per instruction.
// OT = double: the FP64 form; OT = float: the mixed form (FP64 eigen-decomposition and divided differences, FP32
// derivative assembly and FP32 outputs).  PA: powers with one exponent per matrix (alv[i], read by its lane; a
// matrix whose exponent is not finite gets NaN results).  PACK (Voigt input): the packed outputs -- F the 6 Voigt
// values, J the 21 distinct values of dF_ab/dA_cd (Js), H the 126 distinct Voigt second derivatives (6 rows of 21) --
// written as they are staged, without the layout tables.  Used by the homogeneous kernels (smf_kernels.hip, RunRows)
// and the heterogeneous batch kernel (batch.hip, ListRows).
#pragma once
#include "smf_device.hpp"
#include "tile_rows.hpp"

namespace mf {

constexpr int SA = 27;             // phase-A tile per matrix: F (6) + Js (21)
constexpr int SB = 19;             // V (9) + D2 (10)
constexpr int SH = 127;            // staged H values per matrix (6 rows x 21, +1: odd stride)
constexpr int HPP_DEFAULT = 10;    // matrices per phase-B pass (6 lanes each)

// LDS of one tile, in bytes.  Not SEP: one region used in turn by the FP64 input block, the phase-A tile (F, Js),
// then the V / D2 block (offset 0) with the H tile of HPP matrices after it (offset VB).  SEP: the V / D2 block at
// offset VS, after a region used in turn by the input block, the phase-A tile and the H tile (offset 0).
template <bool MAT, int ORDER, class OT, int HPP = HPP_DEFAULT, bool SEP = false>
struct SmfLds {
  static constexpr int NIN = MAT ? 9 : 6;
  static constexpr int IN = WAVE * NIN * 8;
  static constexpr int TA = WAVE * SA * (int)sizeof(OT);
  static constexpr int VBLK = ORDER >= 2 ? ((WAVE * SB * (int)sizeof(OT) + 15) / 16) * 16 : 0;
  static constexpr int HT = ORDER >= 2 ? HPP * SH * (int)sizeof(OT) : 0;
  static constexpr int VB = SEP ? 0 : VBLK;                               // H tile offset
  static constexpr int TB = SEP ? HT : VBLK + HT;
  static constexpr int R1 = IN > TA ? (IN > TB ? IN : TB) : (TA > TB ? TA : TB);
  static constexpr int VS = SEP ? ((R1 + 15) / 16) * 16 : 0;             // V / D2 block offset
  static constexpr int BYTES = SEP ? VS + VBLK : R1;
};
constexpr int smf_nj(bool mat) { return mat ? 81 : 36; }
constexpr int smf_nh(bool mat) { return mat ? 729 : 216; }

// lookup tables of the output layouts (the encoding of copy_out_lut), filled by the wave: J element e reads Js at
// 6 + j_src of its matrix's phase-A tile (stride SA) with weight 2^1 where j_src's weight is 2; H element e reads
// h_src of its matrix's H tile (stride SH) with weight 2^0, 2^-1 or 2^-2.  N + LUT_PAD entries: e >= N stands for
// element e - N of the next matrix
template <bool MAT, int ORDER, class OT>
MF_DEV void smf_luts(short* lutJ, short* lutH, int lane) {
  constexpr int NJ = smf_nj(MAT), NH = smf_nh(MAT);
  if (ORDER >= 1)
    for (int e = lane; e < NJ + LUT_PAD; e += WAVE) {
      double w;
      const int s = 6 + j_src(MAT, e % NJ, w) + (e >= NJ ? SA : 0);
      lutJ[e] = lut_entry<OT>(s, w == 2.0 ? 1 : 0);
    }
  if (ORDER >= 2)
    for (int e = lane; e < NH + LUT_PAD; e += WAVE) {
      double w;
      const int s = h_src(MAT, e % NH, w) + (e >= NH ? SH : 0);
      lutH[e] = lut_entry<OT>(s, w == 1.0 ? 0 : (w == 0.5 ? -1 : -2));
    }
}

// one tile of nm <= 64 matrices; lds: SmfLds<MAT, ORDER, OT, HPP, SEP>::BYTES bytes, 16-byte aligned; the LUTs
// filled (smf_luts) and visible to the wave; lane = threadIdx.x; VEC: elements per lane and store of J and H
// (copy_out, copy_out_lut).  Ends with the tile's outputs written and a barrier (the LDS may be reused).
// pre: a transform of the inputs first (pre_inputs in tile_rows.hpp; matrix layout only)
template <int KIND, bool MAT, int ORDER, class OT, int VEC, bool PA, class R, int HPP = HPP_DEFAULT,
          bool SEP = false, bool PACK = false, class Pre = NoPre>
MF_DEV void smf_tile(const R& rows, int nm, const double* __restrict__ x, OT* __restrict__ F, OT* __restrict__ J,
                     OT* __restrict__ H, double al, const double* __restrict__ alv, const double* __restrict__ c2,
                     unsigned char* lds, const short* lutJ, const short* lutH, int lane, const Pre& pre = Pre()) {
  using L = SmfLds<MAT, ORDER, OT, HPP, SEP>;
  static_assert(!(PACK && MAT), "packed outputs take Voigt input");
  constexpr int NIN = L::NIN, NF = MAT ? 9 : 6, NJ = PACK ? 21 : NF * NF, NH = PACK ? 126 : NF * NF * NF;
  double* xin = reinterpret_cast<double*>(lds);
  OT* t1 = reinterpret_cast<OT*>(lds);             // phase-A tile
  OT* tb = reinterpret_cast<OT*>(lds + L::VS);     // V / D2
  OT* th = reinterpret_cast<OT*>(lds + L::VB);     // H tile

  // ---- input, coalesced through LDS (lanes past the end take the identity)
  load_in<NIN>(xin, x, rows, nm, lane);
  __syncthreads();
  double v[6];
  if (MAT) {
    double a[9];
#pragma unroll
    for (int k = 0; k < 9; k++) a[k] = lane < nm ? xin[9 * lane + k] : (k % 4 == 0 ? 1.0 : 0.0);
    pre_inputs(pre, a, xin, rows, nm, lane);
    matrix_to_voigt(a, v);
  } else {
#pragma unroll
    for (int k = 0; k < 6; k++) v[k] = lane < nm ? xin[6 * lane + k] : (k < 3 ? 1.0 : 0.0);
  }
  bool alok = true;                            // (PA: a non-finite exponent computes with 1 and stores NaN)
  if (PA) {
    al = lane < nm ? alv[rows.m(lane < nm ? lane : 0)] : 1.0;
    alok = fabs(al) <= 1.7976931348623157e308;
    al = alok ? al : 1.0;
  }
  const double nanf = alok ? 1.0 : dnan();
  __syncthreads();

  // ---- phase A
  OT vd[ORDER >= 2 ? SB : 1];                    // V, D2 of this lane's matrix, staged after the F / J stores
  {
    SmfLane<KIND, OT> s;
    double Fv[6];
    s.eigen(v, al, Fv);
#pragma unroll
    for (int k = 0; k < 6; k++) t1[lane * SA + k] = (OT)(Fv[k] * nanf);
    if (ORDER >= 1) {
      OT Js[21];
      s.jacobian(al, Js);
#pragma unroll
      for (int k = 0; k < 21; k++) t1[lane * SA + 6 + k] = Js[k] * (OT)nanf;
    }
    if (ORDER >= 2) {
      s.template second_dd<PA, (sizeof(OT) == 4 ? 20 : SMF_NTERMS)>(al, c2);
#pragma unroll
      for (int k = 0; k < 9; k++) vd[k] = (OT)s.V[k / 3][k % 3];
#pragma unroll
      for (int k = 0; k < 10; k++) vd[9 + k] = (OT)(s.D2[k] * nanf);
      if (SEP) {
#pragma unroll
        for (int k = 0; k < SB; k++) tb[lane * SB + k] = vd[k];
      }
    }
  }
  __syncthreads();
  if constexpr (R::RUN) {
    copy_out<NF, 1, R_IN>(F, rows, NF * nm, lane, [&](int j, int e) { return t1[j * SA + f_src(MAT, e)]; });
  } else {                                     // (rows listed: runs of consecutive rows as blocks)
    for_row_runs<R_IN>(rows, nm, lane, [&](int j0, int j1) {
      copy_run<NF>(F + rows.m(j0) * NF, (j1 - j0) * NF, lane,
                   [&](int j, int e) { return t1[(j0 + j) * SA + f_src(MAT, e)]; });
    });
  }
  if (ORDER >= 1) {
    if constexpr (PACK)
      copy_out<NJ, VEC, R_J>(J, rows, NJ * nm, lane, [&](int j, int e) { return t1[j * SA + 6 + e]; });
    else copy_out_lut<NJ, SA, VEC, R_J>(J, rows, NJ * nm, lane, t1, lutJ);
  }
  if (ORDER >= 2) {
    if (!SEP) {
      __syncthreads();                         // (the phase-A tile fully read)
#pragma unroll
      for (int k = 0; k < SB; k++) tb[lane * SB + k] = vd[k];
    }
    // ---- phase B: lane = (matrix jj of the pass, Voigt row i)
    const int jj = lane / 6, ir = lane - 6 * jj;
    for (int p0 = 0; p0 < nm; p0 += HPP) {
      const int cnt = nm - p0 < HPP ? nm - p0 : HPP;
      __syncthreads();                         // (V / D2 staged; the previous H tile fully read)
      if (jj < cnt) {
        const OT* sb = tb + (p0 + jj) * SB;
        const SmfHess<OT> hs(sb, sb + 9);
        OT h[21];
        hs.h_row(ir, h);
#pragma unroll
        for (int k = 0; k < 21; k++) th[jj * SH + ir * 21 + k] = h[k];
      }
      __syncthreads();
      const R sub = shift_rows(rows, p0);
      if constexpr (PACK)
        copy_out<NH, VEC, R_H>(H, sub, NH * cnt, lane, [&](int j, int e) { return th[j * SH + e]; });
      else copy_out_lut<NH, SH, VEC, R_H>(H, sub, NH * cnt, lane, th, lutH);
    }
  }
  __syncthreads();                             // (the LDS is free for the next tile)
}

}  // namespace mf
