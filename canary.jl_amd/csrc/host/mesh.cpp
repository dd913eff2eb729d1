// Brick mesh generation, Hilbert-curve element ordering and face
// connectivity on one rank, plus the volume-to-face index maps.
//
// Reference behaviour followed (src/mesh.jl):
//   linearpartition :13            hilbertcode :35-88 (Skilling's
//   AxestoTranspose, then the bit transpose into d words, H[0] most
//   significant)   centroidtocode :107-141 (sequential corner sum / nvert,
//   bounding box, floor((2^64-1)*c) evaluated exactly)   brickmesh :259-317
//   getpartition/partition :402-580 (on one rank: elements sorted by their
//   d-word code, then by original index)   vertsortandorder :595-719
//   connectmesh :751-956 (one rank: sort face-vertex tuples, pair equal
//   neighbours, unmatched faces point to themselves)   mappings :973-1001
// Arrays are column-major with 1-based values, exactly as in the reference.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../../include/canary_dg.h"

struct canary_mesh {
  int d = 0;
  int64_t nelem = 0;
  std::vector<int64_t> elemtovert;   // (nvert, nelem)
  std::vector<double> elemtocoord;   // (d, nvert, nelem)
  std::vector<int64_t> elemtobndy;   // (nface, nelem)
  std::vector<std::vector<int64_t>> faceconnections;  // [e, f, verts...]
  bool connected = false;
  std::vector<int64_t> elemtoelem, elemtoface, elemtoordr;  // (nface, nelem)
};

namespace {

// Corner-based face masks: corners (0-based) of face f in Cartesian order.
std::vector<int> corner_fmask(int d) {
  const int nvert = 1 << d, nface = 2 * d, nfv = 1 << (d - 1);
  std::vector<int> fm(nfv * nface);
  for (int f = 0; f < nface; ++f) {
    int fd = f / 2, side = f % 2, c = 0;
    for (int v = 0; v < nvert; ++v)
      if (((v >> fd) & 1) == side) fm[c++ + nfv * f] = v;
  }
  return fm;
}

// floor((2^64 - 1) * c) for 0 <= c <= 1, exactly.
uint64_t scaled_floor(double c) {
  if (!(c > 0.0)) return 0;
  int ex;
  double f = std::frexp(c, &ex);                      // c = f * 2^ex, f in [0.5, 1)
  uint64_t m = (uint64_t)std::ldexp(f, 53);           // exact 53-bit integer
  int shift = ex - 53;                                // c = m * 2^shift
  unsigned __int128 prod = (unsigned __int128)m * (unsigned __int128)UINT64_MAX;
  if (shift >= 0) return (uint64_t)(prod << shift);   // only c == 1 reaches here
  if (-shift >= 128) return 0;
  return (uint64_t)(prod >> (-shift));
}

bool lex_less(const uint64_t* a, const uint64_t* b, int n) {
  for (int i = 0; i < n; ++i) {
    if (a[i] < b[i]) return true;
    if (a[i] > b[i]) return false;
  }
  return false;
}

}  // namespace

extern "C" int canary_hilbertcode(int n, const uint64_t* Y, int bits, uint64_t* H) {
  if (n < 1 || bits < 1 || bits > 64) return CANARY_ERR_ARG;
  std::vector<uint64_t> X(Y, Y + n);
  const uint64_t M = (uint64_t)1 << (bits - 1);
  uint64_t Q = M;
  for (int j = 1; j <= bits - 1; ++j) {
    uint64_t P = Q - 1;
    for (int i = 0; i < n; ++i) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        uint64_t t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
    Q >>= 1;
  }
  for (int i = 1; i < n; ++i) X[i] ^= X[i - 1];
  uint64_t t = 0;
  Q = M;
  for (int j = 1; j <= bits - 1; ++j) {
    if (X[n - 1] & Q) t ^= Q - 1;
    Q >>= 1;
  }
  for (int i = 0; i < n; ++i) X[i] ^= t;
  for (int i = 0; i < n; ++i) H[i] = 0;
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < bits; ++j) {
      int k = i * bits + j;
      uint64_t bit = (X[n - 1 - (k % n)] >> (k / n)) & 1;
      H[n - 1 - i] |= (bit << j);
    }
  }
  return 0;
}

extern "C" int canary_centroidtocode(int d, int nvert, int64_t nelem,
                                     const double* etc, uint64_t* code) {
  if (d < 1 || nvert < 1 || nelem < 0) return CANARY_ERR_ARG;
  std::vector<double> cen((size_t)d * nelem);
  for (int64_t e = 0; e < nelem; ++e)
    for (int j = 0; j < d; ++j) {
      double s = 0.0;
      for (int v = 0; v < nvert; ++v) s += etc[j + (size_t)d * (v + (size_t)nvert * e)];
      cen[j + (size_t)d * e] = s / nvert;
    }
  std::vector<double> cmin(d, INFINITY), cmax(d, -INFINITY), csize(d);
  for (int64_t e = 0; e < nelem; ++e)
    for (int j = 0; j < d; ++j) {
      cmin[j] = std::min(cmin[j], cen[j + (size_t)d * e]);
      cmax[j] = std::max(cmax[j], cen[j + (size_t)d * e]);
    }
  bool allzero = true;
  for (int j = 0; j < d; ++j) {
    csize[j] = cmax[j] - cmin[j];
    if (csize[j] != 0.0) allzero = false;
  }
  if (allzero) {
    for (int j = 0; j < d; ++j) csize[j] = 1.0;
  } else {
    for (int j = 0; j < d; ++j)
      if (csize[j] == 0.0) csize[j] = *std::max_element(csize.begin(), csize.end());
  }
  std::vector<uint64_t> X(d);
  for (int64_t e = 0; e < nelem; ++e) {
    for (int j = 0; j < d; ++j) X[j] = scaled_floor((cen[j + (size_t)d * e] - cmin[j]) / csize[j]);
    canary_hilbertcode(d, X.data(), 64, code + (size_t)d * e);
  }
  return 0;
}

extern "C" int canary_vertsortandorder(int n, const int64_t* v, int64_t* out, int* order) {
  auto mmf = [](int64_t& x, int64_t& y) {
    if (y < x) { std::swap(x, y); return true; }
    return false;
  };
  if (n == 1) {
    out[0] = v[0];
    *order = 1;
    return 0;
  }
  if (n == 2) {
    int64_t a = v[0], b = v[1];
    bool s1 = mmf(a, b);
    out[0] = a; out[1] = b;
    *order = s1 ? 2 : 1;
    return 0;
  }
  if (n == 3) {
    int64_t a = v[0], b = v[1], c = v[2];
    bool s1 = mmf(b, c), s2 = mmf(a, c), s3 = mmf(a, b);
    int o = 0;
    if (!s1 && !s2 && !s3) o = 1;
    else if (!s1 && s2 && s3) o = 2;
    else if (s1 && !s2 && s3) o = 3;
    else if (!s1 && !s2 && s3) o = 4;
    else if (s1 && s2 && s3) o = 5;
    else if (s1 && !s2 && !s3) o = 6;
    out[0] = a; out[1] = b; out[2] = c;
    *order = o;
    return o ? 0 : CANARY_ERR_ORDERING;
  }
  if (n == 4) {
    int64_t a = v[0], b = v[1], c = v[2], dd = v[3];
    bool s1 = mmf(a, b), s2 = mmf(c, dd), s3 = mmf(a, c), s4 = mmf(b, dd), s5 = mmf(b, c);
    int o = 0;
    if (!s1 && !s2 && !s3 && !s4 && !s5) o = 1;
    else if (!s1 && !s2 && !s3 && !s4 && s5) o = 2;
    else if (s1 && !s2 && !s3 && !s4 && !s5) o = 3;
    else if (!s1 && !s2 && s3 && s4 && s5) o = 4;
    else if (s1 && s2 && !s3 && !s4 && s5) o = 5;
    else if (!s1 && !s2 && s3 && s4 && !s5) o = 6;
    else if (s1 && s2 && s3 && s4 && s5) o = 7;
    else if (s1 && s2 && s3 && s4 && !s5) o = 8;
    out[0] = a; out[1] = b; out[2] = c; out[3] = dd;
    *order = o;
    return o ? 0 : CANARY_ERR_ORDERING;
  }
  return CANARY_ERR_ARG;
}

extern "C" canary_mesh* canary_brickmesh(int d, const int64_t* ncoord, const double* coords,
                                         const int* periodic, int64_t part, int64_t numparts,
                                         const int64_t* boundary, int* status) {
  auto fail = [&](int code) -> canary_mesh* {
    if (status) *status = code;
    return nullptr;
  };
  if (d < 1 || d > 3 || numparts < 1 || part < 1 || part > numparts) return fail(CANARY_ERR_ARG);
  for (int j = 0; j < d; ++j)
    if (ncoord[j] < 2) return fail(CANARY_ERR_ARG);
  const int nvert = 1 << d, nface = 2 * d, nfv = 1 << (d - 1);
  std::vector<const double*> x(d);
  size_t off = 0;
  std::vector<int64_t> nel(d), vstride(d), estride(d);
  int64_t ntot = 1, vs = 1;
  for (int j = 0; j < d; ++j) {
    x[j] = coords + off;
    off += ncoord[j];
    nel[j] = ncoord[j] - 1;
    estride[j] = ntot;
    vstride[j] = vs;
    ntot *= nel[j];
    vs *= ncoord[j];
  }
  // linearpartition(n, p, np) = div((p-1)*n, np)+1 : div(p*n, np)
  int64_t lo = ((part - 1) * ntot) / numparts, hi = (part * ntot) / numparts;
  canary_mesh* m = new canary_mesh();
  m->d = d;
  m->nelem = hi - lo;
  m->elemtovert.assign((size_t)nvert * m->nelem, 0);
  m->elemtocoord.assign((size_t)d * nvert * m->nelem, 0.0);
  m->elemtobndy.assign((size_t)nface * m->nelem, 0);
  std::vector<int> fm = corner_fmask(d);
  auto bnd = [&](int side, int dim) -> int64_t {
    return boundary ? boundary[side + 2 * dim] : 1;
  };
  std::vector<int64_t> ec(d), vc(d);
  for (int64_t el = 0; el < m->nelem; ++el) {
    int64_t g = lo + el;  // 0-based global Cartesian element index, x fastest
    for (int j = 0; j < d; ++j) ec[j] = (g / estride[j]) % nel[j];
    for (int v = 0; v < nvert; ++v) {
      int64_t vid = 0;
      for (int j = 0; j < d; ++j) {
        vc[j] = ec[j] + ((v >> j) & 1);
        vid += vc[j] * vstride[j];
        m->elemtocoord[j + (size_t)d * (v + (size_t)nvert * el)] = x[j][vc[j]];
      }
      m->elemtovert[v + (size_t)nvert * el] = vid + 1;
    }
    for (int i = 0; i < d; ++i) {
      if (!periodic[i] && ec[i] == 0) m->elemtobndy[2 * i + (size_t)nface * el] = bnd(0, i);
      if (!periodic[i] && ec[i] == nel[i] - 1) m->elemtobndy[2 * i + 1 + (size_t)nface * el] = bnd(1, i);
    }
    for (int i = 0; i < d; ++i) {
      if (periodic[i] && ec[i] == nel[i] - 1) {
        std::vector<int64_t> fc;
        fc.push_back(el + 1);
        fc.push_back(2 * (i + 1));
        for (int q = 0; q < nfv; ++q) {
          int v = fm[q + nfv * (2 * i)];
          int64_t vid = 0;
          for (int j = 0; j < d; ++j) {
            int64_t cj = (j == i) ? ((v >> j) & 1) : ec[j] + ((v >> j) & 1);
            vid += cj * vstride[j];
          }
          fc.push_back(vid + 1);
        }
        m->faceconnections.push_back(fc);
      }
    }
  }
  if (status) *status = 0;
  return m;
}

extern "C" canary_mesh* canary_mesh_from_arrays(int d, int64_t nelem, const int64_t* etv,
                                                const double* etc, const int64_t* etb,
                                                int64_t nfc, const int64_t* fcs, int* status) {
  if (d < 1 || d > 3 || nelem < 0 || nfc < 0) {
    if (status) *status = CANARY_ERR_ARG;
    return nullptr;
  }
  const int nvert = 1 << d, nface = 2 * d, w = 2 + (1 << (d - 1));
  canary_mesh* m = new canary_mesh();
  m->d = d;
  m->nelem = nelem;
  m->elemtovert.assign(etv, etv + (size_t)nvert * nelem);
  m->elemtocoord.assign(etc, etc + (size_t)d * nvert * nelem);
  m->elemtobndy.assign(etb, etb + (size_t)nface * nelem);
  for (int64_t k = 0; k < nfc; ++k) m->faceconnections.emplace_back(fcs + w * k, fcs + w * (k + 1));
  if (status) *status = 0;
  return m;
}

extern "C" int canary_mesh_partition(canary_mesh* m) {
  if (!m || m->connected) return CANARY_ERR_ARG;
  const int d = m->d, nvert = 1 << d, nface = 2 * d, nfv = 1 << (d - 1);
  const int64_t ne = m->nelem;
  std::vector<int64_t> e2fc((size_t)nfv * nface * ne, 0);
  for (auto& fc : m->faceconnections)
    for (int q = 0; q < nfv; ++q) e2fc[q + (size_t)nfv * ((fc[1] - 1) + (size_t)nface * (fc[0] - 1))] = fc[2 + q];
  auto sort_by_code = [&](std::vector<int64_t>& order) {
    std::vector<uint64_t> code((size_t)d * ne);
    canary_centroidtocode(d, nvert, ne, m->elemtocoord.data(), code.data());
    order.resize(ne);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
      const uint64_t* ca = &code[(size_t)d * a];
      const uint64_t* cb = &code[(size_t)d * b];
      if (lex_less(ca, cb, d)) return true;
      if (lex_less(cb, ca, d)) return false;
      return a < b;
    });
  };
  auto permute = [&](const std::vector<int64_t>& order) {
    std::vector<int64_t> etv(m->elemtovert.size()), etb(m->elemtobndy.size()), efc(e2fc.size());
    std::vector<double> etc(m->elemtocoord.size());
    for (int64_t k = 0; k < ne; ++k) {
      int64_t o = order[k];
      std::memcpy(&etv[(size_t)nvert * k], &m->elemtovert[(size_t)nvert * o], sizeof(int64_t) * nvert);
      std::memcpy(&etc[(size_t)d * nvert * k], &m->elemtocoord[(size_t)d * nvert * o], sizeof(double) * d * nvert);
      std::memcpy(&etb[(size_t)nface * k], &m->elemtobndy[(size_t)nface * o], sizeof(int64_t) * nface);
      std::memcpy(&efc[(size_t)nfv * nface * k], &e2fc[(size_t)nfv * nface * o], sizeof(int64_t) * nfv * nface);
    }
    m->elemtovert.swap(etv);
    m->elemtocoord.swap(etc);
    m->elemtobndy.swap(etb);
    e2fc.swap(efc);
  };
  std::vector<int64_t> order;
  sort_by_code(order);  // getpartition: sort columns [code; elem; rank]
  permute(order);
  sort_by_code(order);  // re-sort of the received elements by [code; index]
  permute(order);
  m->faceconnections.clear();
  for (int64_t e = 0; e < ne; ++e)
    for (int f = 0; f < nface; ++f)
      if (e2fc[(size_t)nfv * (f + (size_t)nface * e)] > 0) {
        std::vector<int64_t> fc = {e + 1, f + 1};
        for (int q = 0; q < nfv; ++q) fc.push_back(e2fc[q + (size_t)nfv * (f + (size_t)nface * e)]);
        m->faceconnections.push_back(fc);
      }
  return 0;
}

extern "C" int canary_mesh_connect(canary_mesh* m) {
  if (!m) return CANARY_ERR_ARG;
  const int d = m->d, nvert = 1 << d, nface = 2 * d, nfv = 1 << (d - 1);
  const int64_t ne = m->nelem, ncol = ne * nface;
  std::vector<int> fm = corner_fmask(d);
  const int W = nfv + 8;
  enum { MR = 0, ME, MF, MO, NR, NE, NF, NO };
  std::vector<int64_t> A((size_t)W * ncol);
  std::vector<int64_t> v(nfv), sv(nfv);
  for (int64_t e = 0; e < ne; ++e)
    for (int f = 0; f < nface; ++f) {
      int64_t j = e * nface + f;
      for (int q = 0; q < nfv; ++q) v[q] = m->elemtovert[fm[q + nfv * f] + (size_t)nvert * e];
      int o = 0;
      int rc = canary_vertsortandorder(nfv, v.data(), sv.data(), &o);
      if (rc) return rc;
      int64_t* col = &A[(size_t)W * j];
      for (int q = 0; q < nfv; ++q) col[q] = sv[q];
      col[nfv + MR] = 0;
      col[nfv + ME] = e + 1;
      col[nfv + MF] = f + 1;
      col[nfv + MO] = o;
      col[nfv + NR] = col[nfv + NE] = col[nfv + NF] = col[nfv + NO] = INT64_MAX;
    }
  for (auto& fc : m->faceconnections) {
    int64_t j = (fc[0] - 1) * nface + (fc[1] - 1);
    int o = 0;
    int rc = canary_vertsortandorder(nfv, &fc[2], sv.data(), &o);
    if (rc) return rc;
    int64_t* col = &A[(size_t)W * j];
    for (int q = 0; q < nfv; ++q) col[q] = sv[q];
    col[nfv + MO] = o;
  }
  std::vector<int64_t> perm(ncol);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](int64_t a, int64_t b) {
    const int64_t* ca = &A[(size_t)W * a];
    const int64_t* cb = &A[(size_t)W * b];
    for (int q = 0; q < nfv; ++q) {
      if (ca[q] < cb[q]) return true;
      if (ca[q] > cb[q]) return false;
    }
    return false;
  });
  auto same = [&](int64_t a, int64_t b) {
    return std::memcmp(&A[(size_t)W * a], &A[(size_t)W * b], sizeof(int64_t) * nfv) == 0;
  };
  for (int64_t k = 0; k < ncol;) {
    int64_t* ca = &A[(size_t)W * perm[k]];
    if (k + 1 < ncol && same(perm[k], perm[k + 1])) {
      int64_t* cb = &A[(size_t)W * perm[k + 1]];
      for (int q = 0; q < 4; ++q) {
        ca[nfv + NR + q] = cb[nfv + MR + q];
        cb[nfv + NR + q] = ca[nfv + MR + q];
      }
      k += 2;
    } else {
      for (int q = 0; q < 4; ++q) ca[nfv + NR + q] = ca[nfv + MR + q];
      k += 1;
    }
  }
  m->elemtoelem.resize((size_t)nface * ne);
  m->elemtoface.resize((size_t)nface * ne);
  m->elemtoordr.assign((size_t)nface * ne, 1);
  for (int64_t e = 0; e < ne; ++e)
    for (int f = 0; f < nface; ++f) {
      m->elemtoelem[f + (size_t)nface * e] = e + 1;
      m->elemtoface[f + (size_t)nface * e] = f + 1;
    }
  for (int64_t j = 0; j < ncol; ++j) {
    const int64_t* c = &A[(size_t)W * j];
    int64_t me = c[nfv + ME], mf = c[nfv + MF], mo = c[nfv + MO];
    int64_t nee = c[nfv + NE], nf = c[nfv + NF], no = c[nfv + NO];
    size_t at = (mf - 1) + (size_t)nface * (me - 1);
    m->elemtoelem[at] = nee;
    m->elemtoface[at] = nf;
    if (d == 2) {
      m->elemtoordr[at] = (no == mo) ? 1 : 2;
    } else {
      if (no != 1 || mo != 1) return CANARY_ERR_ORIENTATION;
      m->elemtoordr[at] = 1;
    }
  }
  m->connected = true;
  return 0;
}

extern "C" int64_t canary_mesh_size(const canary_mesh* m, int what) {
  if (!m) return -1;
  switch (what) {
    case 0: return m->nelem;
    case 1: return m->d;
    case 2: return (int64_t)m->faceconnections.size();
    case 3: return m->connected ? 1 : 0;
  }
  return -1;
}

extern "C" int canary_mesh_get(const canary_mesh* m, const char* name, void* out) {
  if (!m || !name || !out) return CANARY_ERR_ARG;
  auto cp = [&](const void* src, size_t bytes) {
    std::memcpy(out, src, bytes);
    return 0;
  };
  if (!std::strcmp(name, "elemtovert")) return cp(m->elemtovert.data(), m->elemtovert.size() * 8);
  if (!std::strcmp(name, "elemtocoord")) return cp(m->elemtocoord.data(), m->elemtocoord.size() * 8);
  if (!std::strcmp(name, "elemtobndy")) return cp(m->elemtobndy.data(), m->elemtobndy.size() * 8);
  if (!std::strcmp(name, "faceconnections")) {
    int64_t* o = (int64_t*)out;
    size_t w = 2 + ((size_t)1 << (m->d - 1));
    for (size_t k = 0; k < m->faceconnections.size(); ++k)
      std::memcpy(o + w * k, m->faceconnections[k].data(), w * 8);
    return 0;
  }
  if (!m->connected) return CANARY_ERR_ARG;
  if (!std::strcmp(name, "elemtoelem")) return cp(m->elemtoelem.data(), m->elemtoelem.size() * 8);
  if (!std::strcmp(name, "elemtoface")) return cp(m->elemtoface.data(), m->elemtoface.size() * 8);
  if (!std::strcmp(name, "elemtoordr")) return cp(m->elemtoordr.data(), m->elemtoordr.size() * 8);
  return CANARY_ERR_ARG;
}

extern "C" void canary_mesh_free(canary_mesh* m) { delete m; }

extern "C" int canary_mappings(int N, int d, int64_t nelem, const int64_t* elemtoelem,
                               const int64_t* elemtoface, const int64_t* elemtoordr,
                               int64_t* vmapM, int64_t* vmapP) {
  if (N < 1 || d < 1 || d > 3 || nelem < 0) return CANARY_ERR_ARG;
  const int Nq = N + 1, nface = 2 * d;
  int Np = 1, Nfp = 1;
  for (int j = 0; j < d; ++j) Np *= Nq;
  Nfp = Np / Nq;
  // fmask[:, f]: volume nodes with index fe(f) in dimension fd(f), remaining
  // dimensions in Cartesian order (lower dimensions fastest).
  std::vector<int64_t> fmask((size_t)Nfp * nface);
  for (int f = 0; f < nface; ++f) {
    int fd = f / 2, fe = (f % 2) * N;
    int c = 0;
    for (int n = 0; n < Np; ++n) {
      int idx = n;
      for (int j = 0; j < fd; ++j) idx /= Nq;
      if (idx % Nq == fe) fmask[c++ + (size_t)Nfp * f] = n + 1;
    }
  }
  for (int64_t e1 = 0; e1 < nelem; ++e1)
    for (int f1 = 0; f1 < nface; ++f1) {
      size_t at = f1 + (size_t)nface * e1;
      int64_t e2 = elemtoelem[at], f2 = elemtoface[at], o2 = elemtoordr[at];
      if (o2 != 1) return CANARY_ERR_ORIENTATION;
      for (int n = 0; n < Nfp; ++n) {
        vmapM[n + (size_t)Nfp * at] = (int64_t)Np * e1 + fmask[n + (size_t)Nfp * f1];
        vmapP[n + (size_t)Nfp * at] = (int64_t)Np * (e2 - 1) + fmask[n + (size_t)Nfp * (f2 - 1)];
      }
    }
  return 0;
}
