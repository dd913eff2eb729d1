// Generic CDNA4 kernels of the nodal-DG right-hand side and LSRK update.
//
// Templated on the physics policy (phys.h), the number of 1-D nodes NQ = N+1
// and EPW, the number of elements one workgroup owns.  Threads map to
// (element slot, volume node) for volume work and to (element slot, face,
// face node) for face work; a workgroup is sized to a multiple of the
// 64-lane wavefront.
//
// Reference kernels these replace (behaviour, not structure):
//   knl_volumerhs!      examples/3d_kernels/euler3d_set2c.jl:366-486
//   knl_facerhs!        examples/3d_kernels/euler3d_set2c.jl:490-620
//   knl_updatesolution! examples/3d_kernels/euler3d_set2c.jl:624-639
//   (2D: euler2d_set2c.jl:318-524; SWE: shallow_water.jl:303-490)
//
// Evaluation order.  Each per-node result is accumulated in the order of
// the reference CPU kernels (euler3d_set2c.jl:144-358): rhs, then the
// buoyancy source, then the xi sweep n = 1..Nq, the eta sweep, the zeta
// sweep; then the face contributions in ascending face number; then the
// update Q + ((rkb*dt)*rhs)*MJI and rhs*rka.  With -ffp-contract=off and
// the correctly rounded pow of crpow.h (the oracle's pow is correctly
// rounded too) the exact-order kernels give the oracle's bits.
//
// Face contributions are scattered into a staged copy of the element's rhs
// in LDS in `DIM` rounds, faces (2d, 2d+1) per round: the two faces of one
// direction share no node, so the owner computes without atomics, and edge
// and corner nodes receive their face terms in ascending face order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "phys.h"
#include "mfma.h"

namespace canary {

constexpr int round64(int x) { return (x + 63) / 64 * 64; }
constexpr int ipow(int a, int b) { return b == 0 ? 1 : a * ipow(a, b - 1); }

template <class Phys, int NQ>
struct Shape {
  static constexpr int DIM = Phys::DIM;
  static constexpr int Np = ipow(NQ, DIM);
  static constexpr int Nfp = ipow(NQ, DIM - 1);
  static constexpr int NFACE = 2 * DIM;
  static constexpr int NFN = NFACE * Nfp;
};

// Elements per workgroup: smallest EPW whose round64(EPW*Np) keeps >= 90% of
// the lanes busy, else the best utilisation with at most 256 threads.
template <int Np>
constexpr int default_epw() {
  int best = 1;
  double bu = 0.0;
  for (int e = 1; e * Np <= 256 || e == 1; ++e) {
    double u = double(e * Np) / double(round64(e * Np));
    if (u >= 0.9) return e;
    if (u > bu + 1e-12) {
      bu = u;
      best = e;
    }
    if (e * Np > 256) break;
  }
  return best;
}

// Device-side error record of this translation unit:
// [code, face (1-based), element (1-based), offending value].
static __device__ unsigned long long g_status[4];

__device__ __forceinline__ void record_error(unsigned long long code, long long face,
                                             long long elem, long long value) {
  if (atomicCAS(&g_status[0], 0ull, code) == 0ull) {
    g_status[1] = (unsigned long long)face;
    g_status[2] = (unsigned long long)elem;
    g_status[3] = (unsigned long long)value;
  }
}

// Split a 1-based global node id into (element, node), 32-bit when it fits.
template <int Np>
__device__ __forceinline__ void split_id(int64_t id1, int64_t& elem, int& node) {
  const uint64_t id = (uint64_t)(id1 - 1);
  if ((id >> 32) == 0) {
    const uint32_t i32 = (uint32_t)id;
    elem = (int64_t)(i32 / (uint32_t)Np);
    node = (int)(i32 % (uint32_t)Np);
  } else {
    elem = (int64_t)(id / (uint64_t)Np);
    node = (int)(id % (uint64_t)Np);
  }
}

// XCD-aware block order: workgroups are dispatched round-robin over the 8
// XCDs, so give XCD x the contiguous (Hilbert-ordered) chunk of element
// blocks [x*chunk, (x+1)*chunk); face neighbours then share an L2.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t chunk) {
  return (b & 7) * chunk + (b >> 3);
}

// The same over a launch of 8*chunk workgroups walked backwards when
// `reverse` is set.  Stages that ping-pong Q alternate the direction (the
// caller derives it from the buffer order, or the stage parity), so each
// stage starts on the elements the previous one touched last, whose lines
// the memory-side cache still holds (3D N=4, 32^3: 3.8% faster per stage).
// Each element's result does not depend on the order.
__device__ __forceinline__ int64_t stage_block(int64_t b, int64_t chunk, int reverse) {
  return xcd_block(reverse ? 8 * chunk - 1 - b : b, chunk);
}

// ------------------------------------------------------------------ sweeps
// r[s] += sum_m D[m, i] F[m, j, k, s] (xi), then eta, then zeta, with the
// reference CPU order.  sF[d] points at the flux of direction d,
// laid out [s][node].
template <int DIM, int NQ, int NSR, class Real>
__device__ __forceinline__ void volume_sweeps(Real* r, const Real* sD, const Real* sF0,
                                              const Real* sF1, const Real* sF2, int n) {
  constexpr int Np = ipow(NQ, DIM);
  const int i = n % NQ;
  const int j = (n / NQ) % NQ;
  const int k = DIM == 3 ? n / (NQ * NQ) : 0;
  const int base_xi = n - i;                      // (0, j, k)
  const int base_eta = n - j * NQ;                // (i, 0, k)
  const int base_zeta = n - k * NQ * NQ;          // (i, j, 0)
#pragma unroll
  for (int s = 0; s < NSR; ++s) {
    Real a = r[s];
#pragma unroll
    for (int m = 0; m < NQ; ++m) a = a + sD[m + NQ * i] * sF0[s * Np + base_xi + m];
#pragma unroll
    for (int m = 0; m < NQ; ++m) a = a + sD[m + NQ * j] * sF1[s * Np + base_eta + m * NQ];
    if (DIM == 3) {
#pragma unroll
      for (int m = 0; m < NQ; ++m) a = a + sD[m + NQ * k] * sF2[s * Np + base_zeta + m * NQ * NQ];
    }
    r[s] = a;
  }
}

// r[s] += sum_m D[m, i_dir] F[..m..][s] along direction DIR only (same per-state
// accumulation order as volume_sweeps when called for DIR = 0, 1, 2 in turn).
template <int DIM, int NQ, int NSR, class Real, int DIR>
__device__ __forceinline__ void volume_sweep_dir(Real* r, const Real* sD, const Real* sF, int n) {
  constexpr int Np = ipow(NQ, DIM);
  constexpr int STRIDE = DIR == 0 ? 1 : (DIR == 1 ? NQ : NQ * NQ);
  const int idx = DIR == 0 ? n % NQ : (DIR == 1 ? (n / NQ) % NQ : n / (NQ * NQ));
  const int base = n - idx * STRIDE;
#pragma unroll
  for (int s = 0; s < NSR; ++s) {
    Real a = r[s];
#pragma unroll
    for (int m = 0; m < NQ; ++m) a = a + sD[m + NQ * idx] * sF[s * Np + base + m * STRIDE];
    r[s] = a;
  }
}

// ------------------------------------------------------------ face terms
// A face item is one face node `fi` (= f*Nfp + fn) of element e.  Its work
// is split in three steps so that a kernel can issue the loads early and
// consume them later: face_meta (normal, sMJ, maps, bc), face_gather (the
// P-side state from Q) and face_finish (wall state, pressure, Rusanov flux,
// times sMJ -- the amount the reference subtracts from rhs).
template <class Phys>
struct FaceItem {
  using Real = typename Phys::Real;
  Real nrm[Phys::DIM];
  Real sMJ;
  Real qP[Phys::NS];
  Real yP;  // height of the P node (equation sets whose pressure needs it)
  int64_t eP;
  int64_t idM, idP, bcode;  // map and bc entries as loaded (face_load)
  int vidM;  // -1: no contribution (an error was recorded)
  int vidP;
  int bc;
};

// Barrier for threads that exchange data through LDS only: waits for LDS
// traffic (lgkmcnt) but leaves global loads in flight across the barrier.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Volume node (0-based) of face node fn of face f: the canonical face mask
// of src/mesh.jl:978-982 -- dimension f/2 held at its first (even f) or last
// (odd f) index, the other dimensions in column-major order.  vmapM[:,f,e]
// is always Np*(e-1) + this + 1 (src/mesh.jl:994).
template <int DIM, int NQ>
__device__ __forceinline__ int canon_vid(int f, int fn) {
  constexpr int N = NQ - 1;
  const int fe = (f & 1) * N;
  if (DIM == 2) return (f >> 1) == 0 ? fe + NQ * fn : fn + NQ * fe;
  const int a = fn % NQ, b = fn / NQ;
  switch (f >> 1) {
    case 0: return fe + NQ * a + NQ * NQ * b;
    case 1: return a + NQ * fe + NQ * NQ * b;
    default: return a + NQ * b + NQ * NQ * fe;
  }
}

// face_meta in two halves: face_load issues the loads (no use of their
// values, so a kernel can issue more loads behind them), face_resolve checks
// the maps and bc and splits the node ids.
//
// A null vmapM stands for the canonical face mask (the only vmapM that
// src/mesh.jl builds): the M-side id is computed, and the map load re-reads
// the vmapP entry instead (a cache hit, no HBM bytes), so both forms run one
// instruction sequence in one kernel.
template <class Phys, int NQ>
__device__ __forceinline__ void face_load(FaceItem<Phys>& it,
                                          const typename Phys::Real* __restrict__ sgeo,
                                          const int64_t* __restrict__ vmapM,
                                          const int64_t* __restrict__ vmapP,
                                          const int64_t* __restrict__ bndy, int64_t e, int fi) {
  using S = Shape<Phys, NQ>;
  constexpr int DIM = Phys::DIM, NSG = Phys::NSGEO;
  const int f = fi / S::Nfp;
  const size_t fo = (size_t)fi + (size_t)S::NFN * (size_t)e;
  const typename Phys::Real* sg = sgeo + (size_t)NSG * fo;
#pragma unroll
  for (int c = 0; c < DIM; ++c) it.nrm[c] = sg[c];
  it.sMJ = sg[DIM];
  const bool canon = vmapM == nullptr;
  const int64_t rawM = (canon ? vmapP : vmapM)[fo];
  it.idP = vmapP[fo];
  it.idM = canon ? (int64_t)S::Np * e + canon_vid<DIM, NQ>(f, fi - f * S::Nfp) + 1 : rawM;
  it.bcode = bndy[f + S::NFACE * e];
}

// The maps are split and selected without branching on them (a load whose
// value is first used under a branch is moved under it by the compiler, and
// then waiting for it waits for every load issued before); errors are
// recorded in the reference's order of checks: vmapM, then vmapP of an
// interior face, then the bc code.
template <class Phys, int NQ>
__device__ __forceinline__ void face_resolve(FaceItem<Phys>& it, int64_t e, int fi,
                                             int64_t ntotal) {
  using S = Shape<Phys, NQ>;
  constexpr int Np = S::Np;
  const int f = fi / S::Nfp;
  const int64_t idM = it.idM, idP = it.idP, bc = it.bcode;
  const int64_t nodes = ntotal * (int64_t)Np;
  const bool okM = idM >= 1 && idM <= nodes;
  const bool okP = idP >= 1 && idP <= nodes;
  const bool interior = bc == 0;
  const bool good = okM && (bc == 1 || (interior && okP));
  int64_t eP, eM;
  int vP, vM;
  split_id<Np>(okP ? idP : 1, eP, vP);
  split_id<Np>(okM ? idM : 1, eM, vM);
  it.bc = (int)bc;
  it.eP = good && interior ? eP : 0;
  it.vidP = good && interior ? vP : 0;
  it.vidM = good ? vM : -1;  // -1: no contribution
  if (!good) {
    if (!okM)
      record_error(5 /*CANARY_ERR_VMAP*/, f + 1, e + 1, idM);
    else if (interior)
      record_error(5 /*CANARY_ERR_VMAP*/, f + 1, e + 1, idP);
    else
      record_error(4 /*CANARY_ERR_BC*/, f + 1, e + 1, bc);
  }
}

template <class Phys, int NQ>
__device__ __forceinline__ void face_meta(FaceItem<Phys>& it,
                                          const typename Phys::Real* __restrict__ sgeo,
                                          const int64_t* __restrict__ vmapM,
                                          const int64_t* __restrict__ vmapP,
                                          const int64_t* __restrict__ bndy, int64_t e, int fi,
                                          int64_t ntotal) {
  face_load<Phys, NQ>(it, sgeo, vmapM, vmapP, bndy, e, fi);
  face_resolve<Phys, NQ>(it, e, fi, ntotal);
}

// ---------------------------------------------------- prepared geometry
// The compact arrays canary_prepare_geometry builds once from the reference
// ones (vmapM the canonical face mask, every face in the orientation
// `mappings` supports, MJI = 1 / MJ, elemtobndy 0 or 1 -- all checked
// there):
//   pvgeo (Np, DIM*DIM + 1, nelem): metric terms and MJ (no MJI, no x, y, z)
//   psgeo (DIM + 1, Nfp, NFACE, nelem): normal and sMJ (no vMJI)
//   pface int32 (NFACE, nelem): -1 for a wall face (bc 1), else
//         NFACE * eP + fP, the neighbour element and its face (0-based):
//         vmapP[fn, f, e] = Np * eP + fmask[fn, fP] + 1 (src/mesh.jl:994-995)
template <class Phys>
struct Prepared {
  static constexpr int DIM = Phys::DIM;
  static constexpr int NV = DIM * DIM + 1;  // pvgeo fields
  static constexpr int NSG = DIM + 1;       // psgeo fields
};

// face_load over the prepared arrays: the face code goes to idP and is
// turned into node ids by face_decode_resolve once it has arrived.
template <class Phys, int NQ>
__device__ __forceinline__ void face_load_prep(FaceItem<Phys>& it,
                                               const typename Phys::Real* __restrict__ psgeo,
                                               const int32_t* __restrict__ pface, int64_t e,
                                               int fi) {
  using S = Shape<Phys, NQ>;
  constexpr int DIM = Phys::DIM, NSG = Prepared<Phys>::NSG;
  const int f = fi / S::Nfp;
  const size_t fo = (size_t)fi + (size_t)S::NFN * (size_t)e;
  const typename Phys::Real* sg = psgeo + (size_t)NSG * fo;
#pragma unroll
  for (int c = 0; c < DIM; ++c) it.nrm[c] = sg[c];
  it.sMJ = sg[DIM];
  it.idP = pface[f + S::NFACE * e];
}

// face_decode + face_resolve over the prepared arrays, without building the
// 1-based node ids and splitting them again: the code (an int32, as loaded)
// gives the neighbour and its face directly.  The values are those the two
// steps give (the M side is always in range; a neighbour beyond ntotal is
// the same vmap error, with the same payload).
template <class Phys, int NQ>
__device__ __forceinline__ void face_decode_resolve(FaceItem<Phys>& it, int64_t e, int fi,
                                                    int64_t ntotal) {
  using S = Shape<Phys, NQ>;
  constexpr int DIM = Phys::DIM;
  const int f = fi / S::Nfp, fn = fi - f * S::Nfp;
  const int32_t code = (int32_t)it.idP;
  const bool wall = code < 0;
  const int32_t eP = wall ? 0 : code / S::NFACE;
  const int fP = wall ? f : code - eP * S::NFACE;
  const int vP = canon_vid<DIM, NQ>(fP, fn);
  const bool ok = wall || (int64_t)eP < ntotal;
  it.bc = wall ? 1 : 0;
  it.eP = wall || !ok ? 0 : eP;
  it.vidP = wall || !ok ? 0 : vP;
  it.vidM = ok ? canon_vid<DIM, NQ>(f, fn) : -1;  // -1: no contribution
  if (!ok) record_error(5 /*CANARY_ERR_VMAP*/, f + 1, e + 1, (int64_t)S::Np * eP + vP + 1);
}

// The height y of node v of element e (vgeo field iY), for the equation
// sets whose pressure needs it.
template <class Phys, int NQ>
__device__ __forceinline__ typename Phys::Real node_y(const typename Phys::Real* __restrict__ vgeo,
                                                      int64_t e, int v) {
  if constexpr (NeedsY<Phys>::value) {
    constexpr int Np = Shape<Phys, NQ>::Np;
    return vgeo[(size_t)Phys::NVGEO * Np * (size_t)e + (size_t)Phys::iY * Np + v];
  } else {
    return typename Phys::Real(0);
  }
}

template <class Phys, int NQ>
__device__ __forceinline__ void face_gather(FaceItem<Phys>& it,
                                            const typename Phys::Real* __restrict__ Q,
                                            const typename Phys::Real* __restrict__ vgeo) {
  constexpr int Np = Shape<Phys, NQ>::Np, NS = Phys::NS;
  if (it.vidM >= 0 && it.bc == 0) {
    const typename Phys::Real* qe = Q + (size_t)NS * Np * (size_t)it.eP;
#pragma unroll
    for (int s = 0; s < NS; ++s) it.qP[s] = qe[it.vidP + Np * s];
    it.yP = node_y<Phys, NQ>(vgeo, it.eP, it.vidP);
  }
}

// face_gather with a load on every lane: items without a P-side state
// (invalid, wall, bad map) read node 0 of element e_safe, which is in bounds,
// and leave qP to face_finish.
template <class Phys, int NQ>
__device__ __forceinline__ void face_gather_any(FaceItem<Phys>& it,
                                                const typename Phys::Real* __restrict__ Q,
                                                const typename Phys::Real* __restrict__ vgeo,
                                                bool valid, int64_t e_safe) {
  constexpr int Np = Shape<Phys, NQ>::Np, NS = Phys::NS;
  const bool ok = valid && it.vidM >= 0 && it.bc == 0;
  const typename Phys::Real* qe = Q + (size_t)NS * Np * (size_t)(ok ? it.eP : e_safe);
  const int v = ok ? it.vidP : 0;
  typename Phys::Real g[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) g[s] = qe[v + Np * s];
  const typename Phys::Real y = node_y<Phys, NQ>(vgeo, ok ? it.eP : e_safe, v);
  if (ok) {
#pragma unroll
    for (int s = 0; s < NS; ++s) it.qP[s] = g[s];
    it.yP = y;
  }
}

// M-side state through qm_at(s, node) and pm_at(node, qM); out[NSR].
template <class Phys, class QM, class PMF>
__device__ __forceinline__ int face_finish(FaceItem<Phys>& it, const PhysParams& p, QM qm_at,
                                           PMF pm_at, typename Phys::Real* out) {
  using Real = typename Phys::Real;
  constexpr int NS = Phys::NS;
  if (it.vidM < 0) return -1;
  Real qM[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) qM[s] = qm_at(s, it.vidM);
  const Real PM = pm_at(it.vidM, qM);
  Real PP = 0;
  if (it.bc == 0) {
    if (Phys::B_FROM_M) it.qP[NS - 1] = qm_at(NS - 1, it.vidP);
    PP = Phys::HAS_P ? pressure_at<Phys>(it.qP, it.yP, p) : Real(0);
  } else {
    Phys::reflect(qM, PM, it.nrm, it.qP, PP);
  }
  Phys::face_flux(qM, PM, it.qP, PP, it.nrm, it.sMJ, p, out);
  return it.vidM;
}

// All three steps at once.
template <class Phys, int NQ, class QM, class PMF>
__device__ __forceinline__ int face_contribution(const typename Phys::Real* __restrict__ Q,
                                                 const typename Phys::Real* __restrict__ vgeo,
                                                 const typename Phys::Real* __restrict__ sgeo,
                                                 const int64_t* __restrict__ vmapM,
                                                 const int64_t* __restrict__ vmapP,
                                                 const int64_t* __restrict__ bndy, int64_t e,
                                                 int fi, int64_t ntotal, const PhysParams& p,
                                                 QM qm_at, PMF pm_at, typename Phys::Real* out) {
  FaceItem<Phys> it;
  face_meta<Phys, NQ>(it, sgeo, vmapM, vmapP, bndy, e, fi, ntotal);
  face_gather<Phys, NQ>(it, Q, vgeo);
  return face_finish<Phys>(it, p, qm_at, pm_at, out);
}

// The face terms of volume node n, subtracted from r in ascending face order
// (the order facerhs! applies them: it walks the faces in turn, and a node
// lies on at most one face of each pair 2d, 2d+1): with vmapM the canonical
// face mask, the term of face item f*Nfp + fn lands on node canon_vid(f, fn),
// so the node owner finds its one to DIM items from its own indices.  sC:
// the items' terms [fi][s]; sV: their node ids, NOVID where an item has no
// term (an error was recorded).
template <int DIM, int NQ, int NSR, class Real, class VidT>
__device__ __forceinline__ void face_terms_at_node(Real* r, const Real* sC, const VidT* sV, int n,
                                                   VidT novid) {
  constexpr int Nfp = ipow(NQ, DIM - 1);
  const int i = n % NQ, j = (n / NQ) % NQ, k = DIM == 3 ? n / (NQ * NQ) : 0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const int idx = d == 0 ? i : (d == 1 ? j : k);
    if (idx == 0 || idx == NQ - 1) {
      const int f = 2 * d + (idx == 0 ? 0 : 1);
      const int fn = DIM == 2 ? (d == 0 ? j : i)
                              : (d == 0 ? j + NQ * k : (d == 1 ? i + NQ * k : i + NQ * j));
      const int fi = f * Nfp + fn;
      if (sV[fi] != novid) {
#pragma unroll
        for (int s = 0; s < NSR; ++s) r[s] = r[s] - sC[fi * NSR + s];
      }
    }
  }
}

// ================================================================ kernels
template <class Phys, int NQ, int EPW>
struct VolumeCfg {
  using S = Shape<Phys, NQ>;
  static constexpr int NT = round64(EPW * S::Np);
};

// MFMA: the derivative contractions on the matrix cores (MfmaSweep), in
// place in sF, then r + xi + eta + zeta at each node; else volume_sweeps.
template <class Phys, int NQ, int EPW, bool MFMA = false>
__global__ __launch_bounds__((VolumeCfg<Phys, NQ, EPW>::NT)) void volume_kernel(
    double* __restrict__ rhs, const double* __restrict__ Q, const double* __restrict__ vgeo,
    const double* __restrict__ D, int64_t nreal, int64_t chunk, int reverse, PhysParams p) {
  using S = Shape<Phys, NQ>;
  constexpr int Np = S::Np, DIM = S::DIM, NS = Phys::NS, NSR = Phys::NSR, NV = Phys::NVGEO;
  __shared__ double sD[NQ * NQ];
  __shared__ double sF[DIM][EPW][NSR * Np];
  const int64_t blk = stage_block(blockIdx.x, chunk, reverse);
  if (blk * EPW >= nreal) return;  // whole workgroup leaves together
  const int t = threadIdx.x;
  const int slot = t / Np, n = t - slot * Np;
  const int64_t e = blk * EPW + slot;
  const bool valid = slot < EPW && e < nreal;
  if (t < NQ * NQ) sD[t] = D[t];
  double r[NSR];
  if (valid) {
    double q[NS];
    const double* qe = Q + (size_t)NS * Np * (size_t)e;
#pragma unroll
    for (int s = 0; s < NS; ++s) q[s] = qe[n + Np * s];
    const double* vg = vgeo + (size_t)NV * Np * (size_t)e;
    double met[DIM * DIM];
#pragma unroll
    for (int c = 0; c < DIM * DIM; ++c) met[c] = vg[n + Np * c];
    const double MJ = vg[n + Np * Phys::iMJ];
    const double P = Phys::HAS_P ? pressure_at<Phys>(q, node_y<Phys, NQ>(vgeo, e, n), p) : 0.0;
    double F[DIM][NSR];
    Phys::volume_flux(q, P, met, MJ, p, F);
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      // MFMA: each direction line by line (MfmaSweep)
      const int w = MFMA ? MfmaSweep<DIM, NQ, NSR>::slot_of(d, n) : n;
#pragma unroll
      for (int s = 0; s < NSR; ++s) sF[d][slot][s * Np + w] = F[d][s];
    }
    const double* re = rhs + (size_t)NS * Np * (size_t)e;
#pragma unroll
    for (int s = 0; s < NSR; ++s) r[s] = re[n + Np * s];
    Phys::source(q, MJ, p, r);
  }
  lds_barrier();
  if constexpr (MFMA) {
    const int nv = nreal - blk * EPW < EPW ? (int)(nreal - blk * EPW) : EPW;
    mfma_sweeps<DIM, NQ, NSR, VolumeCfg<Phys, NQ, EPW>::NT, EPW, DIM>(
        &sF[0][0][0], EPW * NSR * Np, NSR * Np, nv, sD);
    lds_barrier();
    if (valid) {
      int w[DIM];
#pragma unroll
      for (int d = 0; d < DIM; ++d) w[d] = MfmaSweep<DIM, NQ, NSR>::slot_of(d, n);
#pragma unroll
      for (int s = 0; s < NSR; ++s) {
        double a = r[s];
#pragma unroll
        for (int d = 0; d < DIM; ++d) a = a + sF[d][slot][s * Np + w[d]];
        r[s] = a;
      }
    }
  } else if (valid) {
    volume_sweeps<DIM, NQ, NSR>(r, sD, sF[0][slot], sF[DIM > 1 ? 1 : 0][slot],
                                sF[DIM > 2 ? 2 : 0][slot], n);
  }
  if (valid) {
    double* re = rhs + (size_t)NS * Np * (size_t)e;
#pragma unroll
    for (int s = 0; s < NSR; ++s) re[n + Np * s] = r[s];
  }
}


template <class Phys, int NQ>
__global__ __launch_bounds__(256) void update_kernel(double* __restrict__ rhs,
                                                     double* __restrict__ Q,
                                                     const double* __restrict__ vgeo,
                                                     int64_t nreal, double rka, double rkb,
                                                     double dt, int advection, int64_t chunk,
                                                     int reverse) {
  using S = Shape<Phys, NQ>;
  constexpr int Np = S::Np, NS = Phys::NS, NSR = Phys::NSR, NV = Phys::NVGEO;
  const int64_t total = nreal * Np;
  const double bdt = rkb * dt;
  {
    const int64_t g = stage_block(blockIdx.x, chunk, reverse) * 256 + threadIdx.x;
    if (g >= total) return;
    const int64_t e = g / Np;
    const int n = (int)(g - e * Np);
    const double MJI = vgeo[n + Np * (Phys::iMJI + (size_t)NV * e)];
    double* qe = Q + (size_t)NS * Np * e;
    double* re = rhs + (size_t)NS * Np * e;
    double uu = 0.0, vv = 0.0;
    if (Phys::HAS_ADVECTION && advection) {
      const double Ht = qe[n + Np * 2] + qe[n + Np * 3];
      uu = qe[n] / Ht;
      vv = qe[n + Np] / Ht;
    }
#pragma unroll
    for (int s = 0; s < NSR; ++s) {
      const double r = re[n + Np * s];
      qe[n + Np * s] = qe[n + Np * s] + bdt * r * MJI;
      re[n + Np * s] = r * rka;
    }
    if (Phys::HAS_ADVECTION && advection) {
      const double Ht = qe[n + Np * 2] + qe[n + Np * 3];
      qe[n] = Ht * uu;
      qe[n + Np] = Ht * vv;
    }
  }
}

// --------------------------------------------------------- fused stage
// One LSRK stage of EPW elements per workgroup: volume term, face terms and
// update in one pass over HBM.  Reads Qin (own + neighbour traces), writes
// Qout; rhs in place.
//
// Schedule (per workgroup):
//   1. each thread issues the loads of its first face item (maps, normals,
//      bc), of D and of its volume node's Q, metric terms and rhs, before
//      it uses any of them;
//   2. the P-side gather of the face item is issued as soon as its map has
//      arrived; volume fluxes -> LDS;
//   3. the xi/eta/zeta sweeps run from LDS while the gathers are in flight
//      (barriers fence LDS only, so loads stay outstanding across them);
//   4. face fluxes -> LDS, taken into the staged rhs by each node's thread
//      in ascending face order (or scattered in DIM rounds when vmapM is not
//      the canonical face mask);
//   5. update: Qout = Q + ((rkb*dt)*rhs)*MJI, rhs *= rka, in LDS;
//   6. the block's Qout and rhs written out in storage order.
template <class Phys, int NQ, int EPW, bool WIDE>
struct FusedCfg {
  // vidM of each face node fits a byte up to Np = 255, else a short
  using VidT = typename std::conditional<(Shape<Phys, NQ>::Np < 256), uint8_t, uint16_t>::type;
  using S = Shape<Phys, NQ>;
  using Real = typename Phys::Real;
  // WIDE: enough threads for every face item of the workgroup in one round
  static constexpr int NVOL = round64(EPW * S::Np);
  static constexpr int NFACEIT = round64(EPW * S::NFN);
  static constexpr int NT = WIDE && NFACEIT > NVOL ? NFACEIT : NVOL;
  static constexpr int FLUX = S::Np;                       // per state, one direction
  static constexpr int POST = S::NFN + S::Np;               // sC + sR per state
  static constexpr int BUF = (FLUX > POST ? FLUX : POST) * Phys::NSR;
  // LDS carve-up (bytes, each piece 16-byte aligned)
  static constexpr int a16(int b) { return (b + 15) / 16 * 16; }
  static constexpr int OFF_D = 0;
  static constexpr int OFF_Q = OFF_D + a16(sizeof(Real) * NQ * NQ);
  static constexpr int OFF_P = OFF_Q + a16(sizeof(Real) * EPW * Phys::NS * S::Np);
  static constexpr int OFF_B = OFF_P + a16(sizeof(Real) * EPW * (Phys::HAS_P ? S::Np : 1));
  static constexpr int OFF_V = OFF_B + a16(sizeof(Real) * EPW * BUF);
  static constexpr int OFF_F = OFF_V + a16(sizeof(VidT) * EPW * S::NFN);  // one flag word
  static constexpr int LDS = OFF_F + 16;
};

// One fused stage of the EPW-element block starting at e0, computed by NT
// threads (NT >= the volume nodes of the block).  `lds` holds
// FusedCfg::LDS bytes.  PF: face items per thread whose loads are issued
// before the volume work (the rest are handled afterwards); PM_LDS: keep
// the nodal pressure in LDS for the M side of the faces (else recompute it).
// VOLUME = false leaves the volume term out (face terms + update only: the
// opt-in facerhs_update, after a separate volumerhs); vgeo is then read for
// MJI alone and D is not read.  UPDATE = false (with VOLUME = false) is
// facerhs! alone: rhs -= face terms, Q untouched, neither vgeo nor Qout read.
template <class Phys, int NQ, int EPW, int NT, int PF, bool PM_LDS, bool VOLUME = true,
          bool UPDATE = true, bool PREP = false, bool MFMA = false>
__device__ __forceinline__ void stage_body(
    char* __restrict__ lds, typename Phys::Real* __restrict__ rhs,
    const typename Phys::Real* __restrict__ Qin, typename Phys::Real* __restrict__ Qout,
    const typename Phys::Real* __restrict__ vgeo, const typename Phys::Real* __restrict__ sgeo,
    const double* __restrict__ D, int64_t e0, int64_t nreal, int64_t ntotal,
    const int64_t* __restrict__ vmapM, const int64_t* __restrict__ vmapP,
    const int64_t* __restrict__ bndy, double rka, double rkb, double dt, const PhysParams& p,
    int advection = 0) {
  using S = Shape<Phys, NQ>;
  using C = FusedCfg<Phys, NQ, EPW, false>;
  using Real = typename Phys::Real;
  constexpr int Np = S::Np, DIM = S::DIM, NS = Phys::NS, NSR = Phys::NSR;
  constexpr int NV = PREP ? Prepared<Phys>::NV : Phys::NVGEO;
  constexpr int NFN = S::NFN, Nfp = S::Nfp;
  constexpr int KF = (EPW * NFN + NT - 1) / NT;
  static_assert(!PREP || (VOLUME && UPDATE), "prepared arrays: the whole stage only");
  static_assert(!NeedsY<Phys>::value || (PM_LDS && !PREP),
                "a pressure that needs y: M side from LDS, reference arrays (which hold y)");
  // prepared arrays: psgeo in `sgeo`, the face codes in `vmapP`'s place
  const int32_t* __restrict__ pface = reinterpret_cast<const int32_t*>(vmapP);
  constexpr int KP = PF < KF ? PF : KF;
  constexpr bool SP = Phys::HAS_P && PM_LDS;
  constexpr int BUF = C::BUF;
  using VidT = typename C::VidT;
  constexpr VidT NOVID = (VidT)~(VidT)0;
  static_assert(NT >= EPW * Np, "one volume node per thread");
  static_assert(!MFMA || (VOLUME && std::is_same<Real, double>::value),
                "matrix-core contractions: FP64 volume term");
  Real* sD = reinterpret_cast<Real*>(lds + C::OFF_D);
  Real* sQ = reinterpret_cast<Real*>(lds + C::OFF_Q);  // [EPW][NS*Np]
  Real* sP = reinterpret_cast<Real*>(lds + C::OFF_P);  // [EPW][Np]
  Real* sB = reinterpret_cast<Real*>(lds + C::OFF_B);  // [EPW][BUF]
  VidT* sV = reinterpret_cast<VidT*>(lds + C::OFF_V);  // [EPW][NFN]
  int* sFlag = reinterpret_cast<int*>(lds + C::OFF_F);  // a non-canonical M-side node seen

  const int t = threadIdx.x;
  const int slot = t / Np, n = t - slot * Np;
  const int64_t e = e0 + slot;
  const bool valid = slot < EPW && e < nreal;

  // 1. issue every load of the block before any of them is used: the face
  //    items' normals, maps and bc (t, t + NT, ...), D, the node's Q, metric
  //    terms and rhs -- one memory round trip instead of one per kind.  The
  //    loads are unconditional (lanes without work read in-bounds element
  //    e0, node 0, and ignore it): a load under a branch makes every later
  //    wait count for the path without it, i.e. wait for everything.
  FaceItem<Phys> fit[KP > 0 ? KP : 1];  // KP = 0: every face item after the volume work
  bool fval[KP > 0 ? KP : 1];
#pragma unroll
  for (int k = 0; k < KP; ++k) {
    const int item = t + k * NT;
    const int sl = item / NFN;
    fval[k] = item < EPW * NFN && e0 + sl < nreal;
    if constexpr (PREP)
      face_load_prep<Phys, NQ>(fit[k], sgeo, pface, fval[k] ? e0 + sl : e0,
                               fval[k] ? item - sl * NFN : 0);
    else
      face_load<Phys, NQ>(fit[k], sgeo, vmapM, vmapP, bndy, fval[k] ? e0 + sl : e0,
                          fval[k] ? item - sl * NFN : 0);
  }
  double dval = 0.0;
  if constexpr (VOLUME) dval = D[t < NQ * NQ ? t : 0];
  // keep the map loads ahead of the node loads: waiting for the maps then
  // leaves the node loads in flight (the wait counter is in issue order)
  __builtin_amdgcn_sched_barrier(0);
  const int64_t el = valid ? e : e0;
  const int nl = valid ? n : 0;
  Real q[NS], met[DIM * DIM], r[NSR];
  const Real* qe = Qin + (size_t)NS * Np * (size_t)el;
#pragma unroll
  for (int s = 0; s < NS; ++s) q[s] = qe[nl + Np * s];
  const Real* vg = (VOLUME || UPDATE) ? vgeo + (size_t)NV * Np * (size_t)el : nullptr;
  // the node's height, for a pressure that needs it (loaded with the rest)
  const Real ynode = NeedsY<Phys>::value ? node_y<Phys, NQ>(vgeo, el, nl) : Real(0);
  Real MJ = 0;
  if constexpr (VOLUME) {
#pragma unroll
    for (int c = 0; c < DIM * DIM; ++c) met[c] = vg[nl + Np * c];
    MJ = vg[nl + Np * Phys::iMJ];
  }
  Real MJI = 0;
  if constexpr (UPDATE && !PREP) MJI = vg[nl + Np * Phys::iMJI];
  const Real* re = rhs + (size_t)NS * Np * (size_t)el;
#pragma unroll
  for (int s = 0; s < NSR; ++s) r[s] = re[nl + Np * s];
  // 2. maps checked as soon as they arrive and the P-side gathers issued
  //    (unconditionally, as above: lanes without an item check the item
  //    they loaded, e0's first face node, which some lane owns anyway);
  //    the gathers stay in flight over the volume work and the sweeps
#pragma unroll
  for (int k = 0; k < KP; ++k) {
    const int item = t + k * NT;
    const int sl = item / NFN;
    if constexpr (PREP)
      face_decode_resolve<Phys, NQ>(fit[k], fval[k] ? e0 + sl : e0,
                                    fval[k] ? item - sl * NFN : 0, ntotal);
    else
      face_resolve<Phys, NQ>(fit[k], fval[k] ? e0 + sl : e0, fval[k] ? item - sl * NFN : 0,
                             ntotal);
    face_gather_any<Phys, NQ>(fit[k], Qin, vgeo, fval[k], e0);
  }

  if (VOLUME && t < NQ * NQ) sD[t] = (Real)dval;
  if (!PREP && t == 0) *sFlag = 0;  // read after the face terms, barriers later
  Real Fv[DIM][NSR];
  if (valid) {
#pragma unroll
    for (int s = 0; s < NS; ++s) sQ[slot * NS * Np + s * Np + n] = q[s];
    const Real P = Phys::HAS_P ? pressure_at<Phys>(q, ynode, p) : Real(0);
    if (SP) sP[slot * Np + n] = P;
    if constexpr (VOLUME) {
      Phys::volume_flux(q, P, met, MJ, p, Fv);
      Phys::source(q, MJ, p, r);
    }
  }
  // 3. sweeps, one direction at a time through one flux buffer
#pragma unroll
  for (int d = 0; d < (VOLUME ? DIM : 0); ++d) {
    if (d > 0) lds_barrier();  // the previous direction's fluxes are consumed
    // MFMA: the direction line by line (MfmaSweep)
    const int w = MFMA ? MfmaSweep<DIM, NQ, NSR>::slot_of(d, n) : n;
    if (valid) {
#pragma unroll
      for (int s = 0; s < NSR; ++s) sB[slot * BUF + s * Np + w] = Fv[d][s];
    }
    lds_barrier();
    if constexpr (MFMA) {
      // the direction's contractions in place in sB, then added by the node
      const int nvb = nreal - e0 < EPW ? (int)(nreal - e0) : EPW;
      mfma_sweeps<DIM, NQ, NSR, NT, EPW, 1>(reinterpret_cast<double*>(sB), 0, BUF, nvb,
                                            reinterpret_cast<const double*>(sD));
      lds_barrier();
      if (valid) {
#pragma unroll
        for (int s = 0; s < NSR; ++s) r[s] = r[s] + sB[slot * BUF + s * Np + w];
      }
    } else if (valid) {
      const Real* b = sB + slot * BUF;
      if (d == 0) volume_sweep_dir<DIM, NQ, NSR, Real, 0>(r, sD, b, n);
      if (d == 1) volume_sweep_dir<DIM, NQ, NSR, Real, (DIM > 1 ? 1 : 0)>(r, sD, b, n);
      if (d == 2) volume_sweep_dir<DIM, NQ, NSR, Real, (DIM > 2 ? 2 : 0)>(r, sD, b, n);
    }
  }
  lds_barrier();  // fluxes consumed; sB is reused below
  if (valid) {
#pragma unroll
    for (int s = 0; s < NSR; ++s) sB[slot * BUF + NFN * NSR + s * Np + n] = r[s];
  }
  // 4. face fluxes: M side from LDS, P side gathered
  bool noncanon = false;  // an item whose M-side node is not the canonical one
  auto pm_of = [&](int sl, int v, const Real* qm) -> Real {
    if (!Phys::HAS_P) return Real(0);
    if constexpr (SP)
      return sP[sl * Np + v];
    else
      return Phys::pressure(qm, p);  // (a pressure that needs y always has SP)
  };
#pragma unroll
  for (int k = 0; k < KP; ++k) {
    if (!fval[k]) continue;
    const int item = t + k * NT;
    const int sl = item / NFN, fi = item - sl * NFN;
    auto qm_at = [&](int s, int v) { return sQ[sl * NS * Np + s * Np + v]; };
    auto pm_at = [&](int v, const Real* qm) { return pm_of(sl, v, qm); };
    Real c[NSR];
    const int vid = face_finish<Phys>(fit[k], p, qm_at, pm_at, c);
    if constexpr (!PREP) noncanon |= vid >= 0 && vid != canon_vid<DIM, NQ>(fi / Nfp, fi % Nfp);
    sV[sl * NFN + fi] = vid < 0 ? NOVID : (VidT)vid;
#pragma unroll
    for (int s = 0; s < NSR; ++s) sB[sl * BUF + fi * NSR + s] = c[s];
  }
  for (int item = t + KP * NT; item < EPW * NFN; item += NT) {
    const int sl = item / NFN, fi = item - sl * NFN;
    if (e0 + sl >= nreal) continue;
    auto qm_at = [&](int s, int v) { return sQ[sl * NS * Np + s * Np + v]; };
    auto pm_at = [&](int v, const Real* qm) { return pm_of(sl, v, qm); };
    Real c[NSR];
    int vid;
    if constexpr (PREP) {
      FaceItem<Phys> it;
      face_load_prep<Phys, NQ>(it, sgeo, pface, e0 + sl, fi);
      face_decode_resolve<Phys, NQ>(it, e0 + sl, fi, ntotal);
      face_gather<Phys, NQ>(it, Qin, vgeo);
      vid = face_finish<Phys>(it, p, qm_at, pm_at, c);
    } else {
      vid = face_contribution<Phys, NQ>(Qin, vgeo, sgeo, vmapM, vmapP, bndy, e0 + sl, fi, ntotal, p,
                                        qm_at, pm_at, c);
      noncanon |= vid >= 0 && vid != canon_vid<DIM, NQ>(fi / Nfp, fi % Nfp);
    }
    sV[sl * NFN + fi] = vid < 0 ? NOVID : (VidT)vid;
#pragma unroll
    for (int s = 0; s < NSR; ++s) sB[sl * BUF + fi * NSR + s] = c[s];
  }
  // The face terms reach rhs (staged in LDS) in ascending face order: with
  // the canonical face mask (always over the prepared arrays; over the
  // reference arrays when every item of the workgroup maps to its canonical
  // node) the node owner gathers its own (face_terms_at_node, no barrier);
  // else the face pairs are scattered round by round.  Same operations in
  // the same order either way.
  if (!PREP && noncanon) *sFlag = 1;
  lds_barrier();
  const bool scatter = !PREP && *sFlag != 0;
  if (scatter) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      for (int it = t; it < EPW * 2 * Nfp; it += NT) {
        const int sl = it / (2 * Nfp), fi = 2 * d * Nfp + (it - sl * 2 * Nfp);
        const VidT v = sV[sl * NFN + fi];
        if (e0 + sl < nreal && v != NOVID) {
          Real* rr = sB + sl * BUF + NFN * NSR;
#pragma unroll
          for (int s = 0; s < NSR; ++s)
            rr[s * Np + v] = rr[s * Np + v] - sB[sl * BUF + fi * NSR + s];
        }
      }
      lds_barrier();
    }
  } else if (!UPDATE && valid) {
    Real rv[NSR];
    Real* rr = sB + slot * BUF + NFN * NSR + n;
#pragma unroll
    for (int s = 0; s < NSR; ++s) rv[s] = rr[s * Np];
    face_terms_at_node<DIM, NQ, NSR>(rv, sB + slot * BUF, sV + slot * NFN, n, NOVID);
#pragma unroll
    for (int s = 0; s < NSR; ++s) rr[s * Np] = rv[s];
  }
  // 5. update: Qout = Q + ((rkb*dt)*rhs)*MJI, rhs *= rka, in LDS, by the
  //    thread that holds the node's MJI (which, unless the face terms were
  //    scattered, first takes its own face terms).  Shallow water with
  //    `advection` keeps the velocity: u = Hu / (h + b) before the update,
  //    Hu = (h + b) u after it (shallow_water.jl:274-297).
  if (UPDATE && valid) {
    // prepared arrays hold no MJI: the reference's is 1 ./ MJ (IEEE division)
    if constexpr (PREP) MJI = Real(1) / MJ;
    const Real bdt = (Real)(rkb * dt);
    const Real ka = (Real)rka;
    Real* qn = sQ + slot * NS * Np + n;
    Real uu = 0, vv = 0;
    if (Phys::HAS_ADVECTION && advection) {
      const Real Ht = qn[Np * 2] + qn[Np * 3];
      uu = qn[0] / Ht;
      vv = qn[Np] / Ht;
    }
    Real rv[NSR];
#pragma unroll
    for (int s = 0; s < NSR; ++s) rv[s] = sB[slot * BUF + NFN * NSR + s * Np + n];
    if (!scatter) face_terms_at_node<DIM, NQ, NSR>(rv, sB + slot * BUF, sV + slot * NFN, n, NOVID);
#pragma unroll
    for (int s = 0; s < NSR; ++s) {
      Real& qq = qn[s * Np];
      const Real rs = rv[s];
      qq = qq + bdt * rs * MJI;
      sB[slot * BUF + NFN * NSR + s * Np + n] = rs * ka;
    }
    if (Phys::HAS_ADVECTION && advection) {
      const Real Ht = qn[Np * 2] + qn[Np * 3];
      qn[0] = Ht * uu;
      qn[Np] = Ht * vv;
    }
  }
  lds_barrier();
  // 6. the block's results leave LDS in storage order: consecutive lanes
  //    write consecutive words, so every wave writes whole lines (a store per
  //    state from the node owners writes Np-word runs that start and end
  //    inside lines; 2D N=4: 7-9% slower).  Non-temporal stores: nothing
  //    reads these lines again in this stage, and keeping them out of L2
  //    leaves room for the Q lines the face gathers reuse (3D N=4: 600 fewer
  //    bytes fetched per element, 3% faster).
  const int nv = nreal - e0 < EPW ? (int)(nreal - e0) : EPW;
  if constexpr (UPDATE) {
    Real* qo = Qout + (size_t)NS * Np * (size_t)e0;
    for (int i = t; i < nv * NS * Np; i += NT) __builtin_nontemporal_store(sQ[i], &qo[i]);
  }
  {
    Real* re = rhs + (size_t)NS * Np * (size_t)e0;
    for (int i = t; i < nv * NSR * Np; i += NT) {
      const int sl = i / (NSR * Np), j = i - sl * (NSR * Np);
      __builtin_nontemporal_store(sB[sl * BUF + NFN * NSR + j], &re[(size_t)sl * NS * Np + j]);
    }
  }
}

template <class Phys, int NQ, int EPW, bool WIDE, int PF, bool PM_LDS, bool MFMA = false>
__global__ __launch_bounds__((FusedCfg<Phys, NQ, EPW, WIDE>::NT)) void stage_kernel(
    double* __restrict__ rhs, const double* __restrict__ Qin, double* __restrict__ Qout,
    const double* __restrict__ vgeo, const double* __restrict__ sgeo,
    const double* __restrict__ D, int64_t nreal, int64_t ntotal,
    const int64_t* __restrict__ vmapM, const int64_t* __restrict__ vmapP,
    const int64_t* __restrict__ bndy, double rka, double rkb, double dt, int64_t chunk,
    int reverse, PhysParams p, int advection) {
  using C = FusedCfg<Phys, NQ, EPW, WIDE>;
  __shared__ __attribute__((aligned(16))) char lds[C::LDS];
  const int64_t blk = stage_block(blockIdx.x, chunk, reverse);
  const int64_t e0 = blk * EPW;
  if (e0 >= nreal) return;  // whole workgroup leaves together
  stage_body<Phys, NQ, EPW, C::NT, PF, PM_LDS, true, true, false, MFMA>(
      lds, rhs, Qin, Qout, vgeo, sgeo, D, e0, nreal, ntotal, vmapM, vmapP, bndy, rka, rkb, dt, p,
      advection);
}

// The fused stage over the prepared arrays (see Prepared): the same
// arithmetic and order, fewer bytes.
template <class Phys, int NQ, int EPW, int PF, bool MFMA = false>
__global__ __launch_bounds__((FusedCfg<Phys, NQ, EPW, false>::NT)) void stage_prep_kernel(
    double* __restrict__ rhs, const double* __restrict__ Qin, double* __restrict__ Qout,
    const double* __restrict__ pvgeo, const double* __restrict__ psgeo,
    const int32_t* __restrict__ pface, const double* __restrict__ D, int64_t nreal,
    int64_t ntotal, double rka, double rkb, double dt, int64_t chunk, int reverse, PhysParams p,
    int advection) {
  using C = FusedCfg<Phys, NQ, EPW, false>;
  __shared__ __attribute__((aligned(16))) char lds[C::LDS];
  const int64_t blk = stage_block(blockIdx.x, chunk, reverse);
  const int64_t e0 = blk * EPW;
  if (e0 >= nreal) return;  // whole workgroup leaves together
  stage_body<Phys, NQ, EPW, C::NT, PF, true, true, true, true, MFMA>(
      lds, rhs, Qin, Qout, pvgeo, psgeo, D, e0, nreal, ntotal, nullptr,
      reinterpret_cast<const int64_t*>(pface), nullptr, rka, rkb, dt, p, advection);
}

// facerhs! alone (rhs -= sMJ * Rusanov flux of Q): the fused stage body with
// its volume term and update switched off, so the face terms get the same
// schedule (every load issued first, nodal Q and pressure staged in LDS for
// the M side, non-temporal rhs stores) and the same bits.  vgeo is read only
// by equation sets whose pressure needs the node height (else it may be
// null).
template <class Phys, int NQ, int EPW>
__global__ __launch_bounds__((FusedCfg<Phys, NQ, EPW, false>::NT)) void face_kernel(
    double* __restrict__ rhs, const double* __restrict__ Q, const double* __restrict__ vgeo,
    const double* __restrict__ sgeo, int64_t nreal, int64_t ntotal,
    const int64_t* __restrict__ vmapM, const int64_t* __restrict__ vmapP,
    const int64_t* __restrict__ bndy, int64_t chunk, int reverse, PhysParams p) {
  using C = FusedCfg<Phys, NQ, EPW, false>;
  __shared__ __attribute__((aligned(16))) char lds[C::LDS];
  const int64_t blk = stage_block(blockIdx.x, chunk, reverse);
  const int64_t e0 = blk * EPW;
  if (e0 >= nreal) return;  // whole workgroup leaves together
  stage_body<Phys, NQ, EPW, C::NT, 1, true, false, false>(lds, rhs, Q, nullptr, vgeo, sgeo,
                                                          nullptr, e0, nreal, ntotal, vmapM,
                                                          vmapP, bndy, 0.0, 0.0, 0.0, p);
}

// facerhs! + updatesolution! of one stage in one pass (the volume term was
// added to rhs by volume_kernel before): Qin -> Qout, rhs in place.
template <class Phys, int NQ, int EPW>
__global__ __launch_bounds__((FusedCfg<Phys, NQ, EPW, false>::NT)) void face_update_kernel(
    double* __restrict__ rhs, const double* __restrict__ Qin, double* __restrict__ Qout,
    const double* __restrict__ vgeo, const double* __restrict__ sgeo, int64_t nreal,
    int64_t ntotal, const int64_t* __restrict__ vmapM, const int64_t* __restrict__ vmapP,
    const int64_t* __restrict__ bndy, double rka, double rkb, double dt, int64_t chunk,
    int reverse, PhysParams p, int advection) {
  using C = FusedCfg<Phys, NQ, EPW, false>;
  __shared__ __attribute__((aligned(16))) char lds[C::LDS];
  const int64_t blk = stage_block(blockIdx.x, chunk, reverse);
  const int64_t e0 = blk * EPW;
  if (e0 >= nreal) return;  // whole workgroup leaves together
  stage_body<Phys, NQ, EPW, C::NT, 1, true, false>(lds, rhs, Qin, Qout, vgeo, sgeo, nullptr, e0,
                                                   nreal, ntotal, vmapM, vmapP, bndy, rka, rkb,
                                                   dt, p, advection);
}

}  // namespace canary
