// Precision form of the set2c Euler stage (2D and 3D): the state is split
// into an FP64 background Q0 = (0, .., 0, rho0, rhotheta0) at rest, kept on
// the GPU, and an FP32 perturbation q' = Q - Q0, and only q' is evolved.
// The right-hand side is the FP32 difference L(Q0 + q') - L(Q0) of the
// reference operator (euler3d_set2c.jl:144-348, euler2d_set2c.jl:140-300):
//   * momentum fluxes rho^-1 U_a U_b + delta_ab P', with the pressure
//     perturbation P' = P0 expm1(kappa log1p(rhotheta' / rhotheta0)) (no
//     cancellation against the large background pressure);
//   * rho and rho*theta fluxes are U and rhotheta U / rho (Q0 is at rest);
//   * buoyancy -MJ rho' g on the vertical momentum (the background part
//     balances the background pressure gradient);
//   * Rusanov flux of the perturbation with the full-state wave speed.
// For a hydrostatic (or uniform) background L(Q0) is roundoff (measured
// below 2e-12 of the buoyancy scale in 3D, exactly 0 for a uniform state
// without gravity), so it is not carried.  Metric terms and normals are
// computed in FP64 and rounded to FP32; the background is kept in FP64 and
// its FP32 rounding (the values the FP32 arithmetic uses) is what the stage
// reads; the LSRK update q' + ((rkb*dt) dq) MJI is evaluated in FP64, with
// MJI = 1 / MJ, and rounded once.
//
// Data (per element contiguous, reference-style column-major; NS = dim + 2,
// NF = 2 dim faces):
//   qp, rhs  float (Np, NS, nelem)          perturbation state / LSRK register
//   bg       double (Np, 3, nelem)          rho0, rhotheta0, P0 (reconstruction)
//   bg32     float (Np, 3, nelem)           the same rounded to FP32 (the stage)
//   vg       float (Np, dim*dim + 1, nelem) metric terms, MJ
//   sg       float4 (Nfp, NF, nelem)        (nx, ny, nz, sMJ), nz = 0 in 2D
//   nbr      int32 (NF, nelem)              face code: -1 wall, else NF eP + fP
// prepare checks what the compact arrays leave out: vmapM is the canonical
// face mask of `mappings` (src/mesh.jl:994), each vmapP column is the
// canonical mask of one neighbour face (src/mesh.jl:995), elemtobndy is 0
// or 1, and MJI is 1 ./ MJ.
// Beside the single-solve stage there is a batched stage over a work table
// of (solve, first element) entries, as in batched.hip.
//
// The mixed variant (MIXED = true; canary_perturb_mixed_*) keeps in FP64
// what the compact one rounds: the metric terms, MJ and MJI (vg64, double
// (Np, dim*dim + 2, nelem)), the normals and sMJ (sg64, 4 doubles per face
// node) and the LSRK register dq (rhs, double).  The state q' and the
// volume and face RHS arithmetic stay FP32: the stage's own RHS L(q') is
// summed in FP32 from zero, then dq = dq + L and q' + ((rkb*dt) dq) MJI are
// formed in FP64.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/canary_dg.h"
#include "dg_kernels.h"

namespace canary {
namespace p32 {

static __device__ unsigned long long g_p32_status[4];

__device__ __forceinline__ void p32_error(unsigned long long code, long long face, long long elem,
                                          long long value) {
  if (atomicCAS(&g_p32_status[0], 0ull, code) == 0ull) {
    g_p32_status[1] = (unsigned long long)face;
    g_p32_status[2] = (unsigned long long)elem;
    g_p32_status[3] = (unsigned long long)value;
  }
}

struct P32Params {
  float gamma, kappa, g;
};

__device__ __forceinline__ P32Params params_of(const canary_phys& c) {
  return P32Params{(float)c.gamma, (float)(c.c_p / c.c_v), (float)c.gravity};
}

template <int DIM>
struct Dims {
  static constexpr int NS = DIM + 2;            // U, V, (W), rho, rhotheta
  static constexpr int RHO = DIM, E = DIM + 1;
  static constexpr int NV = DIM * DIM + 1;      // vg fields: metric terms, MJ
  static constexpr int NVG = DIM == 3 ? 14 : 8; // FP64 vgeo fields
  static constexpr int NSG = DIM == 3 ? 5 : 4;  // FP64 sgeo fields
  static constexpr int NF = 2 * DIM;
};

// canonical volume node of face node (f, fn) (src/mesh.jl:973-1001)
template <int DIM, int NQ>
__device__ __forceinline__ int fmask(int f, int fn) {
  constexpr int N = NQ - 1;
  if (DIM == 2) return (f >> 1) == 0 ? (f & 1) * N + NQ * fn : fn + NQ * ((f & 1) * N);
  const int a = fn % NQ, b = fn / NQ;
  switch (f >> 1) {
    case 0: return (f & 1) * N + NQ * a + NQ * NQ * b;
    case 1: return a + NQ * ((f & 1) * N) + NQ * NQ * b;
    default: return a + NQ * b + NQ * NQ * ((f & 1) * N);
  }
}

__device__ __forceinline__ float pprime(float P0, float rt0, float rtp, float kappa) {
  return P0 * expm1f(kappa * log1pf(rtp / rt0));
}

// ------------------------------------------------------------ prepare
template <int DIM, int NQ, class GT>  // GT: geometry word (float compact, double mixed)
__global__ __launch_bounds__(256) void prepare_kernel(
    int64_t nelem, int64_t ntotal, const double* __restrict__ vgeo, const double* __restrict__ sgeo,
    const int64_t* __restrict__ vmapM, const int64_t* __restrict__ vmapP,
    const int64_t* __restrict__ bndy, const double* __restrict__ Q0, const double* __restrict__ Q,
    double kappa_d, double p0, double R, GT* __restrict__ vg, GT* __restrict__ sg,
    int32_t* __restrict__ nbr, double* __restrict__ bg, float* __restrict__ bg32,
    float* __restrict__ qp) {
  using G = Dims<DIM>;
  constexpr int Np = ipow(NQ, DIM), Nfp = ipow(NQ, DIM - 1), NFN = G::NF * Nfp, NS = G::NS;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t g = tid; g < nelem * Np; g += stride) {
    const int64_t e = g / Np;
    const int n = (int)(g - e * Np);
    const double* v = vgeo + (size_t)G::NVG * Np * e;
    GT* o = vg + (size_t)G::NV * Np * e;
#pragma unroll
    for (int c = 0; c < G::NV; ++c) o[n + Np * c] = (GT)v[n + Np * c];
    const double MJ = v[n + Np * (DIM * DIM)], MJI = v[n + Np * (DIM * DIM + 1)];
    if (!(MJI == 1.0 / MJ)) p32_error(1 /*CANARY_ERR_ARG*/, 0, e + 1, n + 1);
    const double* q0 = Q0 + (size_t)NS * Np * e;
    const double* q = Q + (size_t)NS * Np * e;
    const double rho0 = q0[n + G::RHO * Np], rt0 = q0[n + G::E * Np];
    const double P0 = p0 * crpow::pow_cr(R * rt0 / p0, kappa_d);
    double* b = bg + (size_t)3 * Np * e;
    b[n] = rho0;
    b[n + Np] = rt0;
    b[n + 2 * Np] = P0;
    float* b32 = bg32 + (size_t)3 * Np * e;
    b32[n] = (float)rho0;
    b32[n + Np] = (float)rt0;
    b32[n + 2 * Np] = (float)P0;
    float* qo = qp + (size_t)NS * Np * e;
#pragma unroll
    for (int s = 0; s < NS; ++s) qo[n + Np * s] = (float)(q[n + Np * s] - q0[n + Np * s]);
  }
  for (int64_t g = tid; g < nelem * NFN; g += stride) {
    const int64_t e = g / NFN;
    const int fi = (int)(g - e * NFN);
    const int f = fi / Nfp, fn = fi - f * Nfp;
    const double* s = sgeo + (size_t)G::NSG * g;
    GT* so = sg + (size_t)4 * g;
    so[0] = (GT)s[0];
    so[1] = (GT)s[1];
    so[2] = DIM == 3 ? (GT)s[2] : (GT)0;
    so[3] = (GT)s[DIM];
    const int64_t idM = vmapM[g];
    if (idM != (int64_t)Np * e + fmask<DIM, NQ>(f, fn) + 1) p32_error(5, f + 1, e + 1, idM);
  }
  // faces: wall (-1) or the neighbour element and face of the whole vmapP column
  for (int64_t g = tid; g < nelem * G::NF; g += stride) {
    const int64_t e = g / G::NF;
    const int f = (int)(g - e * G::NF);
    const int64_t bc = bndy[g];
    int32_t code = -1;
    if (bc == 0) {
      const int64_t* vp = vmapP + (size_t)Nfp * g;
      const int64_t id0 = vp[0];
      if (id0 < 1 || id0 > ntotal * Np) {
        p32_error(5 /*CANARY_ERR_VMAP*/, f + 1, e + 1, id0);
      } else {
        const int64_t eP = (id0 - 1) / Np;
        int fP = -1;
        for (int c = 0; c < G::NF && fP < 0; ++c) {
          bool all = true;
          for (int k = 0; k < Nfp && all; ++k)
            all = vp[k] == (int64_t)Np * eP + fmask<DIM, NQ>(c, k) + 1;
          if (all) fP = c;
        }
        if (fP < 0)
          p32_error(9 /*CANARY_ERR_ORIENTATION*/, f + 1, e + 1, id0);
        else if (eP > (int64_t)(0x7fffffff / G::NF) - 1)
          p32_error(11 /*CANARY_ERR_CAPACITY*/, f + 1, e + 1, eP + 1);
        else
          code = (int32_t)(G::NF * eP + fP);
      }
    } else if (bc != 1) {
      p32_error(4 /*CANARY_ERR_BC*/, f + 1, e + 1, bc);
    }
    nbr[g] = code;
  }
}

template <int DIM, int NQ>
__global__ __launch_bounds__(256) void reconstruct_kernel(int64_t nelem, const float* __restrict__ qp,
                                                          const double* __restrict__ bg,
                                                          double* __restrict__ Q) {
  using G = Dims<DIM>;
  constexpr int Np = ipow(NQ, DIM), NS = G::NS;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < nelem * Np;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = g / Np;
    const int n = (int)(g - e * Np);
    const float* q = qp + (size_t)NS * Np * e;
    const double* b = bg + (size_t)3 * Np * e;
    double* o = Q + (size_t)NS * Np * e;
#pragma unroll
    for (int s = 0; s < DIM; ++s) o[n + s * Np] = (double)q[n + s * Np];
    o[n + G::RHO * Np] = b[n] + (double)q[n + G::RHO * Np];
    o[n + G::E * Np] = b[n + Np] + (double)q[n + G::E * Np];
  }
}

// ------------------------------------------------------------- stage
template <int DIM, int NQ, int EPW>
struct Cfg32 {
  static constexpr int NS = Dims<DIM>::NS;
  static constexpr int Np = ipow(NQ, DIM), Nfp = ipow(NQ, DIM - 1), NFN = 2 * DIM * Nfp;
  static constexpr int NT = round64(EPW * Np);
  static constexpr int FLUX = DIM * Np * NS;
  static constexpr int POST = NFN * NS + Np * NS;
  static constexpr int BUF = FLUX > POST ? FLUX : POST;
};

template <int DIM, int NQ, int EPW>
struct Lds32 {
  using C = Cfg32<DIM, NQ, EPW>;
  float D[NQ * NQ];
  float Q[EPW][C::NS * C::Np];  // perturbation
  float Bg[EPW][4 * C::Np];     // rho0, rhotheta0, P0, P'
  float B[EPW][C::BUF];         // fluxes, then face terms + rhs
};

// Face record of the geometry word GT: (nx, ny, nz, sMJ) as floats.
__device__ __forceinline__ void load_face(const float* sg, size_t fo, float nn[4]) {
  const float4 s4 = *reinterpret_cast<const float4*>(sg + 4 * fo);
  nn[0] = s4.x;
  nn[1] = s4.y;
  nn[2] = s4.z;
  nn[3] = s4.w;
}
__device__ __forceinline__ void load_face(const double* sg, size_t fo, float nn[4]) {
  const double2 a = *reinterpret_cast<const double2*>(sg + 4 * fo);
  const double2 b = *reinterpret_cast<const double2*>(sg + 4 * fo + 2);
  nn[0] = (float)a.x;
  nn[1] = (float)a.y;
  nn[2] = (float)b.x;
  nn[3] = (float)b.y;
}

// One stage of the EPW-element block at e0 (NT threads, NT >= EPW * Np).
// GT: geometry word (vg, sg); RT: LSRK register word (rhs).  float/float is
// the compact form; double/double the mixed one.
template <int DIM, int NQ, int EPW, class GT = float, class RT = float>
__device__ __forceinline__ void body(Lds32<DIM, NQ, EPW>& L, RT* __restrict__ rhs,
                                     const float* __restrict__ qin, float* __restrict__ qout,
                                     const float* __restrict__ bg, const GT* __restrict__ vg,
                                     const GT* __restrict__ sg, const int32_t* __restrict__ nbr,
                                     const double* __restrict__ D, int64_t e0, int64_t nreal,
                                     double rka, double bdt, const P32Params& p) {
  constexpr bool REG64 = sizeof(RT) == 8;
  using C = Cfg32<DIM, NQ, EPW>;
  using G = Dims<DIM>;
  constexpr int Np = C::Np, Nfp = C::Nfp, NFN = C::NFN, NT = C::NT, NS = G::NS;
  constexpr int RHO = G::RHO, E = G::E, NF = G::NF;
  const int t = threadIdx.x;
  const int slot = t / Np, n = t - slot * Np;
  const int64_t e = e0 + slot;
  const bool valid = slot < EPW && e < nreal;

  // Every load of the block is issued before any is used (face item maps and
  // normals, D, the node's perturbation, metric terms, background and rhs),
  // unconditionally: lanes without work read element e0, node 0, in bounds,
  // and ignore it (a load under a branch makes later waits count for the
  // path without it, i.e. wait for every load).
  const int fsl = t / NFN, ffi = t - fsl * NFN;
  const bool fvalid = t < EPW * NFN && e0 + fsl < nreal;
  const size_t fo = fvalid ? (size_t)ffi + (size_t)NFN * (size_t)(e0 + fsl) : (size_t)NFN * (size_t)e0;
  float fn4[4];
  load_face(sg, fo, fn4);
  const int32_t code_l = nbr[(fvalid ? ffi / Nfp : 0) + NF * (fvalid ? e0 + fsl : e0)];
  const double dval = D[t < NQ * NQ ? t : 0];
  const int64_t el = valid ? e : e0;
  const int nl = valid ? n : 0;
  float r[NS], q[NS], met[DIM * DIM];
  const float* qe = qin + (size_t)NS * Np * el;
#pragma unroll
  for (int s = 0; s < NS; ++s) q[s] = qe[nl + Np * s];
  const GT* v = vg + (size_t)G::NV * Np * el;
#pragma unroll
  for (int c = 0; c < DIM * DIM; ++c) met[c] = (float)v[nl + Np * c];
  const GT MJg = v[nl + Np * DIM * DIM];
  const float MJ = (float)MJg;
  const float* b = bg + (size_t)3 * Np * el;
  const float b0 = b[nl], b1 = b[nl + Np], b2 = b[nl + 2 * Np];
  const RT* re = rhs + (size_t)NS * Np * el;
  // compact: the register is the FP32 accumulator; mixed: the FP64
  // register waits in rreg and the stage's RHS is summed from zero
  RT rreg[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) rreg[s] = re[nl + Np * s];
#pragma unroll
  for (int s = 0; s < NS; ++s) r[s] = REG64 ? 0.f : (float)rreg[s];
  // the P-side gather of the first face item, as soon as its face code is
  // here: node fmask(fP, k) of element eP for face node k (code NF eP + fP)
  const int fbc = fvalid && code_l >= 0 ? 0 : 1;
  const bool gat = fbc == 0;
  const int64_t eP = gat ? (int64_t)(code_l / NF) : e0;
  const int pv = gat ? fmask<DIM, NQ>(code_l - NF * (code_l / NF), ffi % Nfp) : 0;
  float qP[NS];
  {
    const float* qg = qin + (size_t)NS * Np * (size_t)eP;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const float g = qg[pv + Np * s];
      qP[s] = gat ? g : 0.f;
    }
  }
  if (t < NQ * NQ) L.D[t] = (float)dval;
  if (valid) {
    const float rho0 = b0, rt0 = b1, P0 = b2;
    const float Pp = pprime(P0, rt0, q[E], p.kappa);
#pragma unroll
    for (int s = 0; s < NS; ++s) L.Q[slot][s * Np + n] = q[s];
    L.Bg[slot][n] = rho0;
    L.Bg[slot][Np + n] = rt0;
    L.Bg[slot][2 * Np + n] = P0;
    L.Bg[slot][3 * Np + n] = Pp;
    const float rho = rho0 + q[RHO], rt = rt0 + q[E];
    const float ri = 1.0f / rho;
    // physical perturbation fluxes f[c][s] along coordinate c
    float f[DIM][NS];
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      const float uc = q[c] * ri;
#pragma unroll
      for (int s = 0; s < DIM; ++s) f[c][s] = q[s] * uc + (s == c ? Pp : 0.f);
      f[c][RHO] = q[c];
      f[c][E] = rt * uc;
    }
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        float a = 0.f;
#pragma unroll
        for (int c = 0; c < DIM; ++c) a += met[d + DIM * c] * f[c][s];
        L.B[slot][(d * NS + s) * Np + n] = MJ * a;
      }
    }
    r[DIM - 1] -= MJ * q[RHO] * p.g;
  }
  lds_barrier();
  if (valid) {
    const float* F0 = L.B[slot];
    volume_sweeps<DIM, NQ, NS, float>(r, L.D, F0, F0 + NS * Np, F0 + (DIM > 2 ? 2 : 0) * NS * Np,
                                      n);
  }
  lds_barrier();
  if (valid) {
#pragma unroll
    for (int s = 0; s < NS; ++s) L.B[slot][NFN * NS + s * Np + n] = r[s];
  }
  auto face = [&](int sl, int fi, const float* nn, float* qPl, int bc) {
    const int vid = fmask<DIM, NQ>(fi / Nfp, fi % Nfp);
    float qM[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) qM[s] = L.Q[sl][s * Np + vid];
    const float rho0 = L.Bg[sl][vid], rt0 = L.Bg[sl][Np + vid], P0 = L.Bg[sl][2 * Np + vid];
    const float PpM = L.Bg[sl][3 * Np + vid];
    const float sMJ = nn[3];
    float PpP;
    if (bc == 0) {
      PpP = pprime(P0, rt0, qPl[E], p.kappa);
    } else {
      float un = 0.f;
#pragma unroll
      for (int c = 0; c < DIM; ++c) un += nn[c] * qM[c];
#pragma unroll
      for (int c = 0; c < DIM; ++c) qPl[c] = qM[c] - 2.f * un * nn[c];
      qPl[RHO] = qM[RHO];
      qPl[E] = qM[E];
      PpP = PpM;
    }
    const float rM = rho0 + qM[RHO], rP = rho0 + qPl[RHO];
    const float rtM = rt0 + qM[E], rtP = rt0 + qPl[E];
    float UnM = 0.f, UnP = 0.f;
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      UnM += nn[c] * qM[c];
      UnP += nn[c] * qPl[c];
    }
    const float unM = UnM / rM, unP = UnP / rP;
    const float lM = fabsf(unM) + sqrtf(p.gamma * (P0 + PpM) / rM);
    const float lP = fabsf(unP) + sqrtf(p.gamma * (P0 + PpP) / rP);
    const float l = fmaxf(lM, lP);
    float nfM[NS], nfP[NS];
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      nfM[c] = qM[c] * unM + nn[c] * PpM;
      nfP[c] = qPl[c] * unP + nn[c] * PpP;
    }
    nfM[RHO] = UnM;
    nfP[RHO] = UnP;
    nfM[E] = rtM * unM;
    nfP[E] = rtP * unP;
#pragma unroll
    for (int s = 0; s < NS; ++s)
      L.B[sl][fi * NS + s] = sMJ * (0.5f * (nfM[s] + nfP[s] - l * (qPl[s] - qM[s])));
  };
  if (fvalid) face(fsl, ffi, fn4, qP, fbc);
  for (int item = t + NT; item < EPW * NFN; item += NT) {
    const int sl = item / NFN, fi = item - sl * NFN;
    if (e0 + sl >= nreal) continue;
    const size_t fo = (size_t)fi + (size_t)NFN * (size_t)(e0 + sl);
    float nn[4];
    load_face(sg, fo, nn);
    const int32_t code = nbr[fi / Nfp + NF * (e0 + sl)];
    const int bc = code >= 0 ? 0 : 1;
    float ql[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) ql[s] = 0.f;
    if (bc == 0) {
      const int32_t ep = code / NF;
      const float* qe = qin + (size_t)NS * Np * (size_t)ep;
      const int pn = fmask<DIM, NQ>(code - NF * ep, fi % Nfp);
#pragma unroll
      for (int s = 0; s < NS; ++s) ql[s] = qe[pn + Np * s];
    }
    face(sl, fi, nn, ql, bc);
  }
  lds_barrier();
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    for (int it = t; it < EPW * 2 * Nfp; it += NT) {
      const int sl = it / (2 * Nfp), fi = 2 * d * Nfp + (it - sl * 2 * Nfp);
      if (e0 + sl < nreal) {
        const int v = fmask<DIM, NQ>(fi / Nfp, fi % Nfp);
        float* rr = &L.B[sl][NFN * NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) rr[s * Np + v] -= L.B[sl][fi * NS + s];
      }
    }
    lds_barrier();
  }
  if (valid) {  // non-temporal stores: not read again in this stage
    const double MJI = 1.0 / (double)MJg;  // the reference's MJI is 1 ./ MJ
    float* qo = qout + (size_t)NS * Np * e;
    RT* ro = rhs + (size_t)NS * Np * e;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const float rs = L.B[slot][NFN * NS + s * Np + n];
      if constexpr (REG64) {
        const double dq = (double)rreg[s] + (double)rs;
        __builtin_nontemporal_store((float)((double)q[s] + (bdt * dq) * MJI), &qo[n + Np * s]);
        __builtin_nontemporal_store(rka * dq, &ro[n + Np * s]);
      } else {
        __builtin_nontemporal_store((float)((double)q[s] + (bdt * (double)rs) * MJI),
                                    &qo[n + Np * s]);
        __builtin_nontemporal_store((float)rka * rs, &ro[n + Np * s]);
      }
    }
  }
}

template <int DIM, int NQ>
constexpr int single_epw() {
  return default_epw<ipow(NQ, DIM)>();
}

template <int DIM, int NQ>
__global__ __launch_bounds__((Cfg32<DIM, NQ, single_epw<DIM, NQ>()>::NT)) void stage32_kernel(
    float* __restrict__ rhs, const float* __restrict__ qin, float* __restrict__ qout,
    const float* __restrict__ bg, const float* __restrict__ vg, const float* __restrict__ sg,
    const int32_t* __restrict__ nbr, const double* __restrict__ D, int64_t nreal, float rka,
    double bdt, int64_t chunk, int reverse, P32Params p) {
  constexpr int EPW = single_epw<DIM, NQ>();
  __shared__ Lds32<DIM, NQ, EPW> L;
  const int64_t e0 = stage_block(blockIdx.x, chunk, reverse) * EPW;
  if (e0 >= nreal) return;
  body<DIM, NQ, EPW>(L, rhs, qin, qout, bg, vg, sg, nbr, D, e0, nreal, (double)rka, bdt, p);
}

// the mixed variant: FP64 geometry and LSRK register
template <int DIM, int NQ>
__global__ __launch_bounds__((Cfg32<DIM, NQ, single_epw<DIM, NQ>()>::NT)) void stage_mixed_kernel(
    double* __restrict__ rhs, const float* __restrict__ qin, float* __restrict__ qout,
    const float* __restrict__ bg, const double* __restrict__ vg, const double* __restrict__ sg,
    const int32_t* __restrict__ nbr, const double* __restrict__ D, int64_t nreal, double rka,
    double bdt, int64_t chunk, int reverse, P32Params p) {
  constexpr int EPW = single_epw<DIM, NQ>();
  __shared__ Lds32<DIM, NQ, EPW> L;
  const int64_t e0 = stage_block(blockIdx.x, chunk, reverse) * EPW;
  if (e0 >= nreal) return;
  body<DIM, NQ, EPW, double, double>(L, rhs, qin, qout, bg, vg, sg, nbr, D, e0, nreal, rka, bdt,
                                     p);
}

// batched: one workgroup per work-table entry, entries of one XCD contiguous
template <int DIM, int NQ, int EPW>
__global__ __launch_bounds__((Cfg32<DIM, NQ, EPW>::NT)) void batch32_kernel(
    const canary_p32_solve* __restrict__ solves, const canary_batch_item* __restrict__ items,
    int64_t nitems, int64_t chunk, int parity, double rka, double rkb) {
  __shared__ Lds32<DIM, NQ, EPW> L;
  const int64_t w = stage_block(blockIdx.x, chunk, parity);  // direction alternates per stage
  if (w >= nitems) return;
  const canary_batch_item it = items[w];
  const canary_p32_solve& sv = solves[it.solve];
  if (it.N != NQ - 1 || it.epw != EPW || sv.N != NQ - 1 || sv.dim != DIM) {
    if (threadIdx.x == 0) p32_error(1 /*CANARY_ERR_ARG*/, it.solve, it.e0 + 1, it.N);
    return;
  }
  if (it.e0 < 0 || it.e0 >= sv.nreal) {
    if (threadIdx.x == 0) p32_error(1 /*CANARY_ERR_ARG*/, it.solve, it.e0 + 1, sv.nreal);
    return;
  }
  const P32Params p = params_of(sv.phys);
  body<DIM, NQ, EPW>(L, (float*)sv.rhs, (const float*)sv.q[parity], (float*)sv.q[1 - parity],
                     (const float*)sv.bg32, (const float*)sv.vg, (const float*)sv.sg, sv.nbr,
                     sv.D, it.e0, sv.nreal, rka, rkb * sv.dt, p);
}

int launch_status() { return hipGetLastError() == hipSuccess ? 0 : CANARY_ERR_LAUNCH; }

template <int DIM, int NQ, class GT = float>
int prepare(int64_t nelem, int64_t ntotal, const void* vgeo, const void* sgeo, const int64_t* vmapM,
            const int64_t* vmapP, const int64_t* bndy, const void* Q0, const void* Q,
            const canary_phys* ph, void* vg, void* sg, int32_t* nbr, void* bg, void* bg32,
            void* qp, hipStream_t st) {
  const int64_t work = nelem * ipow(NQ, DIM);
  int64_t blocks = (work + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  prepare_kernel<DIM, NQ, GT><<<dim3((unsigned)blocks), dim3(256), 0, st>>>(
      nelem, ntotal, (const double*)vgeo, (const double*)sgeo, vmapM, vmapP, bndy,
      (const double*)Q0, (const double*)Q, ph->c_p / ph->c_v, ph->p0, ph->R_gas, (GT*)vg,
      (GT*)sg, nbr, (double*)bg, (float*)bg32, (float*)qp);
  return launch_status();
}

template <int DIM, int NQ>
int prepare_mixed(int64_t nelem, int64_t ntotal, const void* vgeo, const void* sgeo,
                  const int64_t* vmapM, const int64_t* vmapP, const int64_t* bndy, const void* Q0,
                  const void* Q, const canary_phys* ph, void* vg, void* sg, int32_t* nbr,
                  void* bg, void* bg32, void* qp, hipStream_t st) {
  return prepare<DIM, NQ, double>(nelem, ntotal, vgeo, sgeo, vmapM, vmapP, bndy, Q0, Q, ph, vg, sg,
                                  nbr, bg, bg32, qp, st);
}

template <int DIM, int NQ>
int stage_mixed(void* rhs, const void* qin, void* qout, const void* bg, const void* vg,
                const void* sg, const int32_t* nbr, const double* D, int64_t nreal, double rka,
                double rkb, double dt, const canary_phys* ph, hipStream_t st) {
  constexpr int EPW = single_epw<DIM, NQ>();
  using C = Cfg32<DIM, NQ, EPW>;
  const int64_t nb = (nreal + EPW - 1) / EPW;
  const int64_t chunk = (nb + 7) / 8;
  P32Params p{(float)ph->gamma, (float)(ph->c_p / ph->c_v), (float)ph->gravity};
  stage_mixed_kernel<DIM, NQ><<<dim3((unsigned)(8 * chunk)), dim3(C::NT), 0, st>>>(
      (double*)rhs, (const float*)qin, (float*)qout, (const float*)bg, (const double*)vg,
      (const double*)sg, nbr, D, nreal, rka, rkb * dt, chunk,
      (uintptr_t)qin > (uintptr_t)qout ? 1 : 0, p);
  return launch_status();
}

template <int DIM, int NQ>
int stage(void* rhs, const void* qin, void* qout, const void* bg, const void* vg, const void* sg,
          const int32_t* nbr, const double* D, int64_t nreal, double rka, double rkb, double dt,
          const canary_phys* ph, hipStream_t st) {
  constexpr int EPW = single_epw<DIM, NQ>();
  using C = Cfg32<DIM, NQ, EPW>;
  const int64_t nb = (nreal + EPW - 1) / EPW;
  const int64_t chunk = (nb + 7) / 8;
  P32Params p{(float)ph->gamma, (float)(ph->c_p / ph->c_v), (float)ph->gravity};
  stage32_kernel<DIM, NQ><<<dim3((unsigned)(8 * chunk)), dim3(C::NT), 0, st>>>(
      (float*)rhs, (const float*)qin, (float*)qout, (const float*)bg, (const float*)vg,
      (const float*)sg, nbr, D, nreal, (float)rka, rkb * dt, chunk,
      (uintptr_t)qin > (uintptr_t)qout ? 1 : 0,  // q ping-pong: direction alternates
      p);
  return launch_status();
}

template <int DIM, int NQ>
int reconstruct(int64_t nelem, const void* qp, const void* bg, void* Q, hipStream_t st) {
  const int64_t work = nelem * ipow(NQ, DIM);
  int64_t blocks = (work + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  reconstruct_kernel<DIM, NQ><<<dim3((unsigned)blocks), dim3(256), 0, st>>>(
      nelem, (const float*)qp, (const double*)bg, (double*)Q);
  return launch_status();
}

// ------------------------------------------------------ batched variants
constexpr int MAXT = 256;

// candidate WHICH of elements per workgroup: 0 = most that fit MAXT
// threads, 1 = about half of that, 2 = the single-solve choice; 0 = none
template <int DIM, int NQ, int WHICH>
constexpr int epw_candidate() {
  constexpr int Np = ipow(NQ, DIM);
  constexpr int emax = MAXT / Np;
  if (WHICH == 0) return emax > 0 ? emax : 1;  // more than MAXT nodes: one element alone
  if (WHICH == 1) return emax >= 2 ? (emax + 1) / 2 : 0;
  if (WHICH == 2) return single_epw<DIM, NQ>() * Np <= MAXT ? single_epw<DIM, NQ>() : 0;
  return 0;
}

template <int DIM, int NQ, int WHICH>
int batch_which(int* epw, int* threads, const canary_p32_solve* solves,
                const canary_batch_item* items, int64_t nitems, int parity, double rka,
                double rkb, hipStream_t st) {
  constexpr int E = epw_candidate<DIM, NQ, WHICH>();
  if constexpr (E < 1) {
    return CANARY_ERR_UNSUPPORTED;
  } else {
    constexpr int NT = Cfg32<DIM, NQ, E>::NT;
    if (epw) {
      *epw = E;
      *threads = NT;
      return 0;
    }
    const int64_t chunk = (nitems + 7) / 8;
    if (chunk > ((int64_t)1 << 27)) return CANARY_ERR_ARG;
    batch32_kernel<DIM, NQ, E><<<dim3((unsigned)(8 * chunk)), dim3(NT), 0, st>>>(
        solves, items, nitems, chunk, parity, rka, rkb);
    return launch_status();
  }
}

template <int DIM, int NQ>
int batch_nq(int which, int want, int* epw, int* threads, const canary_p32_solve* solves,
             const canary_batch_item* items, int64_t nitems, int parity, double rka, double rkb,
             hipStream_t st) {
  if (which >= 0) {
    if (which == 0) return batch_which<DIM, NQ, 0>(epw, threads, 0, 0, 0, 0, 0, 0, 0);
    if (which == 1) return batch_which<DIM, NQ, 1>(epw, threads, 0, 0, 0, 0, 0, 0, 0);
    if (which == 2) return batch_which<DIM, NQ, 2>(epw, threads, 0, 0, 0, 0, 0, 0, 0);
    return CANARY_ERR_UNSUPPORTED;
  }
  constexpr int E0 = epw_candidate<DIM, NQ, 0>();
  constexpr int E1 = epw_candidate<DIM, NQ, 1>();
  constexpr int E2 = epw_candidate<DIM, NQ, 2>();
  if (want == E0 && E0 > 0)
    return batch_which<DIM, NQ, 0>(nullptr, nullptr, solves, items, nitems, parity, rka, rkb, st);
  if (want == E1 && E1 > 0)
    return batch_which<DIM, NQ, 1>(nullptr, nullptr, solves, items, nitems, parity, rka, rkb, st);
  if (want == E2 && E2 > 0)
    return batch_which<DIM, NQ, 2>(nullptr, nullptr, solves, items, nitems, parity, rka, rkb, st);
  return CANARY_ERR_UNSUPPORTED;
}

}  // namespace p32
}  // namespace canary

// N -> NQ dispatch, N = 1..7
#define P32_DISPATCH(DIM_, N, FN, ARGS)                    \
  if ((DIM_) == 3) {                                       \
    switch (N) {                                           \
      case 1: return canary::p32::FN<3, 2> ARGS;           \
      case 2: return canary::p32::FN<3, 3> ARGS;           \
      case 3: return canary::p32::FN<3, 4> ARGS;           \
      case 4: return canary::p32::FN<3, 5> ARGS;           \
      case 5: return canary::p32::FN<3, 6> ARGS;           \
      case 6: return canary::p32::FN<3, 7> ARGS;           \
      case 7: return canary::p32::FN<3, 8> ARGS;           \
    }                                                      \
  } else if ((DIM_) == 2) {                                \
    switch (N) {                                           \
      case 1: return canary::p32::FN<2, 2> ARGS;           \
      case 2: return canary::p32::FN<2, 3> ARGS;           \
      case 3: return canary::p32::FN<2, 4> ARGS;           \
      case 4: return canary::p32::FN<2, 5> ARGS;           \
      case 5: return canary::p32::FN<2, 6> ARGS;           \
      case 6: return canary::p32::FN<2, 7> ARGS;           \
      case 7: return canary::p32::FN<2, 8> ARGS;           \
    }                                                      \
  }                                                        \
  return CANARY_ERR_UNSUPPORTED;

extern "C" int canary_perturb32_prepare(int dim, int N, int64_t nelem, int64_t nelem_total,
                                        const void* vgeo, const void* sgeo, const int64_t* vmapM,
                                        const int64_t* vmapP, const int64_t* elemtobndy,
                                        const void* Q0, const void* Q, const canary_phys* phys,
                                        void* vg32, void* sg32, int32_t* nbr, void* bg,
                                        void* bg32, void* qp, void* stream) {
  if (nelem < 0 || nelem_total < nelem) return CANARY_ERR_ARG;
  if (nelem == 0) return 0;
  if (!vgeo || !sgeo || !vmapM || !vmapP || !elemtobndy || !Q0 || !Q || !vg32 || !sg32 || !nbr ||
      !bg || !bg32 || !qp)
    return CANARY_ERR_ARG;
  canary_phys d;
  if (!phys) {
    canary_phys_default(&d);
    phys = &d;
  }
  hipStream_t st = (hipStream_t)stream;
  P32_DISPATCH(dim, N, prepare, (nelem, nelem_total, vgeo, sgeo, vmapM, vmapP, elemtobndy, Q0, Q,
                                 phys, vg32, sg32, nbr, bg, bg32, qp, st))
}

extern "C" int canary_stage_perturb32(int dim, int N, void* rhs, const void* qin, void* qout,
                                      const void* bg32, const void* vg32, const void* sg32,
                                      const int32_t* nbr, const double* D, int64_t nreal,
                                      double rka, double rkb, double dt, const canary_phys* phys,
                                      void* stream) {
  if (nreal < 0) return CANARY_ERR_ARG;
  if (nreal == 0) return 0;
  if (!rhs || !qin || !qout || !bg32 || !vg32 || !sg32 || !nbr || !D || qin == qout)
    return CANARY_ERR_ARG;
  if (((uintptr_t)sg32 & 15) != 0) return CANARY_ERR_ARG;  // read as float4 records
  canary_phys d;
  if (!phys) {
    canary_phys_default(&d);
    phys = &d;
  }
  hipStream_t st = (hipStream_t)stream;
  P32_DISPATCH(dim, N, stage,
               (rhs, qin, qout, bg32, vg32, sg32, nbr, D, nreal, rka, rkb, dt, phys, st))
}

extern "C" int canary_perturb_mixed_prepare(int dim, int N, int64_t nelem, int64_t nelem_total,
                                           const void* vgeo, const void* sgeo,
                                           const int64_t* vmapM, const int64_t* vmapP,
                                           const int64_t* elemtobndy, const void* Q0,
                                           const void* Q, const canary_phys* phys, void* vg64,
                                           void* sg64, int32_t* nbr, void* bg, void* bg32,
                                           void* qp, void* stream) {
  if (nelem < 0 || nelem_total < nelem) return CANARY_ERR_ARG;
  if (nelem == 0) return 0;
  if (!vgeo || !sgeo || !vmapM || !vmapP || !elemtobndy || !Q0 || !Q || !vg64 || !sg64 || !nbr ||
      !bg || !bg32 || !qp)
    return CANARY_ERR_ARG;
  canary_phys d;
  if (!phys) {
    canary_phys_default(&d);
    phys = &d;
  }
  hipStream_t st = (hipStream_t)stream;
  P32_DISPATCH(dim, N, prepare_mixed, (nelem, nelem_total, vgeo, sgeo, vmapM, vmapP, elemtobndy,
                                       Q0, Q, phys, vg64, sg64, nbr, bg, bg32, qp, st))
}

extern "C" int canary_stage_perturb_mixed(int dim, int N, void* rhs64, const void* qin,
                                          void* qout, const void* bg32, const void* vg64,
                                          const void* sg64, const int32_t* nbr, const double* D,
                                          int64_t nreal, double rka, double rkb, double dt,
                                          const canary_phys* phys, void* stream) {
  if (nreal < 0) return CANARY_ERR_ARG;
  if (nreal == 0) return 0;
  if (!rhs64 || !qin || !qout || !bg32 || !vg64 || !sg64 || !nbr || !D || qin == qout)
    return CANARY_ERR_ARG;
  if (((uintptr_t)sg64 & 15) != 0) return CANARY_ERR_ARG;  // read as double2 pairs
  canary_phys d;
  if (!phys) {
    canary_phys_default(&d);
    phys = &d;
  }
  hipStream_t st = (hipStream_t)stream;
  P32_DISPATCH(dim, N, stage_mixed,
               (rhs64, qin, qout, bg32, vg64, sg64, nbr, D, nreal, rka, rkb, dt, phys, st))
}

extern "C" int canary_perturb32_reconstruct(int dim, int N, int64_t nelem, const void* qp,
                                            const void* bg, void* Q, void* stream) {
  if (nelem < 0) return CANARY_ERR_ARG;
  if (nelem == 0) return 0;
  if (!qp || !bg || !Q) return CANARY_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  P32_DISPATCH(dim, N, reconstruct, (nelem, qp, bg, Q, st))
}

extern "C" int canary_p32_batch_variant(int dim, int N, int which, int* epw, int* threads) {
  if (!epw || !threads || which < 0) return CANARY_ERR_ARG;
  P32_DISPATCH(dim, N, batch_nq,
               (which, 0, epw, threads, nullptr, nullptr, 0, 0, 0.0, 0.0, nullptr))
}

extern "C" int canary_p32_batch_stage(int dim, int N, int epw, const canary_p32_solve* solves,
                                      const canary_batch_item* items, int64_t nitems, int parity,
                                      double rka, double rkb, void* stream) {
  if (nitems < 0 || (parity != 0 && parity != 1) || epw < 1) return CANARY_ERR_ARG;
  if (nitems == 0) return 0;
  if (!solves || !items) return CANARY_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  P32_DISPATCH(dim, N, batch_nq,
               (-1, epw, nullptr, nullptr, solves, items, nitems, parity, rka, rkb, st))
}

extern "C" int canary_k_perturb32_status(int64_t* info, int clear) {
  unsigned long long h[4] = {0, 0, 0, 0};
  if (hipMemcpyFromSymbol(h, HIP_SYMBOL(canary::p32::g_p32_status), sizeof(h), 0,
                          hipMemcpyDeviceToHost) != hipSuccess)
    return CANARY_ERR_LAUNCH;
  if (clear && h[0] != 0) {
    unsigned long long z[4] = {0, 0, 0, 0};
    if (hipMemcpyToSymbol(HIP_SYMBOL(canary::p32::g_p32_status), z, sizeof(z), 0,
                          hipMemcpyHostToDevice) != hipSuccess)
      return CANARY_ERR_LAUNCH;
  }
  if (info)
    for (int i = 0; i < 4; ++i) info[i] = (int64_t)h[i];
  return (int)h[0];
}
