// The derivative contractions of the volume term on the matrix cores.
//
// Along direction d the volume term adds, at every node (i along d, `line`
// the other indices), sum_m D[m + NQ*i] F[s][node(d, line, m)] -- for each
// state s and line a product of the NQ x NQ matrix D^T with the line's NQ
// fluxes.  With NQ <= 8 one v_mfma_f64_16x16x4f64 tile takes NB = 16 / NQ
// of those products side by side:
//
//   A (16 x 16) = NB copies of D^T on the diagonal (zero elsewhere),
//   B (16 x 16): row kb*NQ + m, column c holds the flux at position m of
//                line `kb*16 + c` of the tile,
//   A * B: row rb*NQ + i, column c = the sum for position i of line
//          `rb*16 + c`,
//
// in KS = ceil(NB*NQ / 4) steps of K = 4.  A tile covers 16*NB (state, line)
// columns; the tiles of one direction cover disjoint columns, and each
// tile's result goes back to the LDS words its B operand was read from, so
// the contraction runs in place in the flux buffer.  The buffer of a
// direction is laid out line by line, [s][line][position along the line]
// (word (s*NLINE + line)*NQ + m, slot_of gives a node's), so a lane's B
// operand and result words are a column times NQ plus the position: a few
// integer operations per tile.  Waves take whole tiles (the MFMA runs on a
// whole wave).  The sum over m is done by the matrix core, in its own order
// and rounding, so the result is within the contracted form's tolerance of
// the exact order, not its bits.
//
// Lane layout of the FP64 16x16x4 MFMA (one double per lane for A and B,
// four for the result), lane l = 0..63:
//   A[l % 16][l / 16], B[l / 16][l % 16], result r: row l / 16 + 4*r,
//   column l % 16 (checked on gfx950 with integer-valued operands).
#pragma once
#include <hip/hip_runtime.h>

namespace canary {

typedef double mfma_v4d __attribute__((ext_vector_type(4)));

template <int DIM, int NQ, int NSR>
struct MfmaSweep {
  static_assert(NQ >= 2 && NQ <= 8, "one line's D^T must fit a 16 x 16 tile");
  static constexpr int NB = 16 / NQ;                // line products per tile
  static constexpr int KS = (NB * NQ + 3) / 4;      // K = 4 steps
  static constexpr int NLINE = DIM == 3 ? NQ * NQ : (DIM == 2 ? NQ : 1);
  static constexpr int NCOL = NSR * NLINE;          // (state, line) columns
  static constexpr int COLS = 16 * NB;              // columns per tile
  static constexpr int NTILE = (NCOL + COLS - 1) / COLS;
  static constexpr int Np = NQ * NLINE;

  // Word of node n (within its state's Np words) in direction d's
  // line-by-line layout: (line)*NQ + position along d.
  __device__ static __forceinline__ int slot_of(int d, int n) {
    if (d == 0) return n;  // lines along xi are already contiguous
    const int i = n % NQ;
    if (DIM == 2) return i * NQ + n / NQ;
    if (d == 1) return (i + NQ * (n / (NQ * NQ))) * NQ + (n / NQ) % NQ;
    return (n % (NQ * NQ)) * NQ + n / (NQ * NQ);
  }

  // The tile-invariant part of a lane's work: its A operand of each K step
  // (the block-diagonal D^T) and, for its B operand words and its four
  // result words, the offset from column*NQ and the bound on its column in
  // line product 0 (0: the word is padding).
  struct Lane {
    double a[KS];
    int boff[KS], blim[KS];
    int ooff[4], olim[4];
  };

  __device__ static __forceinline__ void lane_setup(const double* sD, int lane, Lane& L) {
    const int row = lane % 16, rb = row / NQ, i = row - rb * NQ;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int k = 4 * ks + lane / 16, kb = k / NQ, m = k - kb * NQ;
      const bool on = rb == kb && rb < NB;
      const double v = sD[on ? m + NQ * i : 0];
      L.a[ks] = on ? v : 0.0;
      L.boff[ks] = kb * 16 * NQ + m;
      L.blim[ks] = kb < NB ? NCOL - kb * 16 : 0;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int orow = lane / 16 + 4 * r, ob = orow / NQ, oi = orow - ob * NQ;
      L.ooff[r] = ob * 16 * NQ + oi;
      L.olim[r] = ob < NB ? NCOL - ob * 16 : 0;
    }
  }

  // One tile, in place in `buf` (one direction of one element, line by line).
  __device__ static __forceinline__ void tile(double* buf, int t, const Lane& L, int lane) {
    const int c = t * COLS + lane % 16;  // this lane's column in line product 0
    const int cw = c * NQ;
    mfma_v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const bool on = c < L.blim[ks];
      const double v = buf[on ? cw + L.boff[ks] : 0];
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(L.a[ks], on ? v : 0.0, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (c < L.olim[r]) buf[cw + L.ooff[r]] = acc[r];
  }
};

// The tiles of ND directions of the valid elements of a workgroup, spread
// over its NT / 64 waves.  Direction dd of slot sl lives at
// base + dd*dstride + sl*sstride, line by line (slot_of).  Every wave runs
// whole tiles (the loop bounds are the same on all lanes of a wave).
template <int DIM, int NQ, int NSR, int NT, int EPW, int ND>
__device__ __forceinline__ void mfma_sweeps(double* base, int dstride, int sstride, int nvalid,
                                            const double* sD) {
  using M = MfmaSweep<DIM, NQ, NSR>;
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  typename M::Lane L;
  M::lane_setup(sD, lane, L);
  const int nitem = ND * nvalid * M::NTILE;
  for (int it = w; it < nitem; it += NW) {
    const int dd = it / (nvalid * M::NTILE);
    const int rest = it - dd * nvalid * M::NTILE;
    const int sl = rest / M::NTILE, t = rest - sl * M::NTILE;
    M::tile(base + dd * dstride + sl * sstride, t, L, lane);
  }
}

}  // namespace canary
