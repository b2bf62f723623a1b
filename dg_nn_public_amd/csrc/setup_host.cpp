// Host-side set-up: element tables, topology, geometry, P1CG graph, colouring, SFC.
#include "setup_host.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <numeric>

namespace dgmi {
namespace {

constexpr int P = 3;

// barycentric multi-indices (sum P) in the reference's local node order
// (z02/shape_function.py:511-532 tets, :1127-1140 triangles)
const int TET_INDEX[20][4] = {
    {3, 0, 0, 0}, {0, 3, 0, 0}, {0, 0, 3, 0}, {0, 0, 0, 3}, {2, 0, 1, 0},
    {1, 0, 2, 0}, {2, 1, 0, 0}, {1, 2, 0, 0}, {0, 2, 1, 0}, {0, 1, 2, 0},
    {2, 0, 0, 1}, {1, 0, 0, 2}, {0, 2, 0, 1}, {0, 1, 0, 2}, {0, 0, 2, 1},
    {0, 0, 1, 2}, {0, 1, 1, 1}, {1, 1, 1, 0}, {1, 0, 1, 1}, {1, 1, 0, 1}};
const int TRI_INDEX[10][3] = {{3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {2, 1, 0}, {1, 2, 0},
                              {0, 2, 1}, {0, 1, 2}, {1, 0, 2}, {2, 0, 1}, {1, 1, 1}};
// face vertex lists (z02/mesh_init.py:305-320, :43-45)
const int TET_FACE_VERTS[4][3] = {{2, 1, 3}, {0, 2, 3}, {1, 0, 3}, {0, 1, 2}};
const int TRI_FACE_VERTS[3][2] = {{0, 1}, {1, 2}, {2, 0}};
// d(xi)/d(face params) (z02/function_space.py:200-213)
const double TET_DRST[4][3][2] = {{{0, 0}, {0, 1}, {1, 0}},
                                  {{1, 0}, {0, 0}, {0, 1}},
                                  {{0, 1}, {1, 0}, {0, 0}},
                                  {{-1, -1}, {1, 0}, {0, 1}}};
const double TRI_DRST[3][2] = {{-1.0, 1.0}, {0.0, -1.0}, {1.0, 0.0}};

// distinct permutations of a 4-tuple, in the order itertools.permutations yields them
void orbit4(const std::array<double, 4>& v, std::vector<std::array<double, 4>>& out) {
  std::array<int, 4> idx = {0, 1, 2, 3};
  std::vector<std::array<double, 4>> seen;
  do {
    std::array<double, 4> t = {v[idx[0]], v[idx[1]], v[idx[2]], v[idx[3]]};
    bool dup = false;
    for (auto& s : seen)
      if (s == t) { dup = true; break; }
    if (!dup) seen.push_back(t);
  } while (std::next_permutation(idx.begin(), idx.end()));
  out.insert(out.end(), seen.begin(), seen.end());
}

// Keast #7 (z02/shape_function.py:25-64): returns barycentric (ngi, 4) and weights
void tet_rule_24(std::vector<double>& lam, std::vector<double>& w) {
  struct Orb { std::array<double, 4> g; double w; };
  const Orb orbs[4] = {
      {{0.3561913862225449, 0.2146028712591517, 0.2146028712591517, 0.2146028712591517}, 0.0399227502581679},
      {{0.8779781243961660, 0.0406739585346113, 0.0406739585346113, 0.0406739585346113}, 0.0100772110553207},
      {{0.0329863295731731, 0.3223378901422757, 0.3223378901422757, 0.3223378901422757}, 0.0553571815436544},
      {{0.2696723314583159, 0.6030056647916491, 0.0636610018750175, 0.0636610018750175}, 0.0482142857142857}};
  lam.clear();
  w.clear();
  for (auto& o : orbs) {
    std::vector<std::array<double, 4>> pts;
    orbit4(o.g, pts);
    for (auto& p : pts) {
      double l1 = p[0], l2 = p[1], l3 = p[2];
      lam.push_back(l1);
      lam.push_back(l2);
      lam.push_back(l3);
      lam.push_back(1.0 - (l1 + l2 + l3));
      w.push_back(o.w / 6.0);
    }
  }
}

// Strang-Fix #9 (z02/shape_function.py:271-308): barycentric (12, 3), weights sum 1/2
void tri_rule_12(std::vector<double>& pts, std::vector<double>& w) {
  const double a = 0.873821971016996, b = 0.063089014491502;
  const double c = 0.501426509658179, d = 0.249286745170910;
  const double e = 0.636502499121399, f = 0.310352451033785, g = 0.053145049844816;
  std::vector<std::array<double, 2>> xy = {{a, b}, {b, a}, {b, b}, {c, d}, {d, c}, {d, d}};
  std::array<double, 3> efg = {e, f, g};
  std::array<int, 3> idx = {0, 1, 2};
  do {
    xy.push_back({efg[idx[0]], efg[idx[1]]});
  } while (std::next_permutation(idx.begin(), idx.end()));
  pts.clear();
  for (auto& p : xy) {
    pts.push_back(p[0]);
    pts.push_back(p[1]);
    pts.push_back(1.0 - (p[0] + p[1]));
  }
  w.assign(12, 0.0);
  for (int i = 0; i < 3; ++i) w[i] = 0.050844906370207 * 0.5;
  for (int i = 3; i < 6; ++i) w[i] = 0.116786275726379 * 0.5;
  for (int i = 6; i < 12; ++i) w[i] = 0.082851075618374 * 0.5;
}

// 4-point Gauss-Legendre on [0,1] (z02/shape_function.py:392-420)
void line_rule_4(std::vector<double>& pts, std::vector<double>& w) {
  const double a = 0.339981043584856, b = 0.861136311594053;
  const double w1 = 0.652145154862546, w2 = 0.347854845137454;
  const double s[4] = {0.5 - 0.5 * b, 0.5 - 0.5 * a, 0.5 + 0.5 * a, 0.5 + 0.5 * b};
  pts.clear();
  for (int i = 0; i < 4; ++i) {
    pts.push_back(s[i]);
    pts.push_back(1.0 - s[i]);
  }
  w = {w2 / 2.0, w1 / 2.0, w1 / 2.0, w2 / 2.0};
}

// f_k(t) = prod_{j<k} (P t - j)/(j+1) and derivative
inline void factor(int k, double t, double& val, double& der) {
  val = 1.0;
  der = 0.0;
  for (int j = 0; j < k; ++j) {
    double term = (P * t - j) / (j + 1);
    der = der * term + val * (double(P) / (j + 1));
    val = val * term;
  }
}

// phi (nloc) and dphi/dxi (D, nloc) at one barycentric point
void lagrange(const std::vector<int>& index, int nb, const double* lam, double* phi,
              double* dxi, int nloc) {
  int D = nb - 1;
  std::vector<double> vals(nb), ders(nb), dl(nb);
  for (int n = 0; n < nloc; ++n) {
    for (int i = 0; i < nb; ++i) factor(index[n * nb + i], lam[i], vals[i], ders[i]);
    double p = 1.0;
    for (int i = 0; i < nb; ++i) p *= vals[i];
    phi[n] = p;
    for (int i = 0; i < nb; ++i) {
      double prod = ders[i];
      for (int k = 0; k < nb; ++k)
        if (k != i) prod *= vals[k];
      dl[i] = prod;
    }
    for (int d = 0; d < D; ++d) dxi[d * nloc + n] = dl[d] - dl[D];
  }
}

// element barycentrics of the point with face barycentrics p on face f
// (z02/shape_function.py:763-783 tets, :1105-1116 triangles)
void face_lam(int D, int f, const double* p, double* lam) {
  for (int i = 0; i < 4; ++i) lam[i] = 0.0;
  if (D == 3) {
    const double p0 = p[0], p1 = p[1], p2 = p[2];
    const double tab[4][4] = {{0, p1, p0, p2}, {p0, 0, p1, p2}, {p1, p0, 0, p2}, {p0, p1, p2, 0}};
    for (int i = 0; i < 4; ++i) lam[i] = tab[f][i];
  } else {
    const double s0 = p[0], s1 = p[1];
    const double tab[3][3] = {{s0, s1, 0}, {0, s0, s1}, {s1, 0, s0}};
    for (int i = 0; i < 3; ++i) lam[i] = tab[f][i];
  }
}

}  // namespace

ElementTables make_element(int D) {
  ElementTables el;
  el.D = D;
  el.nface = D + 1;
  const int nb = D + 1;
  std::vector<double> vlam, vw, fp, fw;
  if (D == 3) {
    el.nloc = 20;
    el.snloc = 10;
    for (int n = 0; n < 20; ++n)
      for (int i = 0; i < 4; ++i) el.index.push_back(TET_INDEX[n][i]);
    tet_rule_24(vlam, vw);
    tri_rule_12(fp, fw);
  } else {
    el.nloc = 10;
    el.snloc = 4;
    for (int n = 0; n < 10; ++n)
      for (int i = 0; i < 3; ++i) el.index.push_back(TRI_INDEX[n][i]);
    tri_rule_12(vlam, vw);
    line_rule_4(fp, fw);
  }
  const int nloc = el.nloc;
  el.ngi = (int)vw.size();
  el.sngi = (int)fw.size();
  el.weight = vw;
  el.sweight = fw;
  el.n.assign(nloc * el.ngi, 0.0);
  el.nlx.assign(D * nloc * el.ngi, 0.0);
  std::vector<double> phi(nloc), dxi(D * nloc);
  for (int g = 0; g < el.ngi; ++g) {
    lagrange(el.index, nb, &vlam[g * nb], phi.data(), dxi.data(), nloc);
    for (int n = 0; n < nloc; ++n) {
      el.n[n * el.ngi + g] = phi[n];
      for (int d = 0; d < D; ++d) el.nlx[(d * nloc + n) * el.ngi + g] = dxi[d * nloc + n];
    }
  }
  // face points in element barycentrics
  const int nf = el.nface, sg = el.sngi;
  el.sn.assign(nf * nloc * sg, 0.0);
  el.snlx.assign(nf * D * nloc * sg, 0.0);
  for (int f = 0; f < nf; ++f) {
    for (int g = 0; g < sg; ++g) {
      double lam[4];
      face_lam(D, f, &fp[g * D], lam);
      lagrange(el.index, nb, lam, phi.data(), dxi.data(), nloc);
      for (int n = 0; n < nloc; ++n) {
        el.sn[(f * nloc + n) * sg + g] = phi[n];
        for (int d = 0; d < D; ++d)
          el.snlx[((f * D + d) * nloc + n) * sg + g] = dxi[d * nloc + n];
      }
    }
  }
  // alignment: face point g on this side has weights fp[g] on the face's vertex list;
  // the neighbour lists the same vertices reversed and rotated to start at this side's
  // vertex k (alnmt = k).  Row k maps g to the neighbour's matching point.
  el.gi_align.assign(D * sg, 0);
  for (int k = 0; k < D; ++k) {
    int nb_order[3];
    if (D == 3) {
      const int rev[3] = {0, 2, 1};
      int start = 0;
      for (int i = 0; i < 3; ++i)
        if (rev[i] == k) start = i;
      for (int i = 0; i < 3; ++i) nb_order[i] = rev[(start + i) % 3];
    } else {
      nb_order[0] = (k == 0) ? 0 : 1;
      nb_order[1] = (k == 0) ? 1 : 0;
    }
    for (int g = 0; g < sg; ++g) {
      int best = -1;
      double bestd = 1e300;
      for (int h = 0; h < sg; ++h) {
        double d = 0.0;
        for (int i = 0; i < D; ++i) {
          // neighbour point h weight on this-side vertex nb_order[i] is fp[h][i]
          double wn = fp[h * D + i];
          double wt = fp[g * D + nb_order[i]];
          d += std::fabs(wn - wt);
        }
        if (d < bestd) { bestd = d; best = h; }
      }
      el.gi_align[k * sg + g] = best;
    }
  }
  // quadratic interpolation of face gradients: nodes in face barycentrics, their Lagrange
  // basis at the face points, gradients there; checked against the tabulated snlx
  {
    const int fb = D;  // face barycentric coordinates per point
    std::vector<std::array<double, 3>> qn;
    if (D == 3)
      qn = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0.5, 0.5, 0}, {0, 0.5, 0.5}, {0.5, 0, 0.5}};
    else
      qn = {{1, 0, 0}, {0, 1, 0}, {0.5, 0.5, 0}};
    const int nk = (int)qn.size();
    el.nfk = nk;
    auto quad_basis = [&](const double* mu, double* L) {
      if (D == 3) {
        for (int i = 0; i < 3; ++i) L[i] = mu[i] * (2.0 * mu[i] - 1.0);
        L[3] = 4.0 * mu[0] * mu[1];
        L[4] = 4.0 * mu[1] * mu[2];
        L[5] = 4.0 * mu[0] * mu[2];
      } else {
        L[0] = mu[0] * (2.0 * mu[0] - 1.0);
        L[1] = mu[1] * (2.0 * mu[1] - 1.0);
        L[2] = 4.0 * mu[0] * mu[1];
      }
    };
    el.flk.assign(sg * nk, 0.0);
    for (int g = 0; g < sg; ++g) quad_basis(&fp[g * fb], &el.flk[g * nk]);
    el.fgk.assign(nf * nk * D * nloc, 0.0);
    el.fnodes.clear();
    el.fphi.assign(nf * sg * el.snloc, 0.0);
    double err = 0.0, big = 0.0;
    for (int f = 0; f < nf; ++f) {
      for (int k = 0; k < nk; ++k) {
        double lam[4];
        face_lam(D, f, qn[k].data(), lam);
        lagrange(el.index, nb, lam, phi.data(), dxi.data(), nloc);
        for (int d = 0; d < D; ++d)
          for (int n = 0; n < nloc; ++n) el.fgk[((f * nk + k) * D + d) * nloc + n] = dxi[d * nloc + n];
      }
      std::vector<int> on;
      for (int n = 0; n < nloc; ++n) {
        double m = 0.0;
        for (int g = 0; g < sg; ++g) m = std::max(m, std::fabs(el.sn[(f * nloc + n) * sg + g]));
        if (m != 0.0) on.push_back(n);
      }
      if ((int)on.size() != el.snloc) err = INFINITY;
      for (int i = 0; i < (int)on.size() && i < el.snloc; ++i) {
        el.fnodes.push_back(on[i]);
        for (int g = 0; g < sg; ++g)
          el.fphi[(f * sg + g) * el.snloc + i] = el.sn[(f * nloc + on[i]) * sg + g];
      }
      for (int d = 0; d < D; ++d)
        for (int n = 0; n < nloc; ++n)
          for (int g = 0; g < sg; ++g) {
            double v = 0.0;
            for (int k = 0; k < nk; ++k)
              v += el.flk[g * nk + k] * el.fgk[((f * nk + k) * D + d) * nloc + n];
            const double ref = el.snlx[((f * D + d) * nloc + n) * sg + g];
            err = std::max(err, std::fabs(v - ref));
            big = std::max(big, std::fabs(ref));
          }
    }
    el.face_interp_err = big > 0.0 ? err / big : err;
    // face mass matrix and mixed face matrix of the node-form face terms
    const int sn = el.snloc;
    el.fmf.assign(nf * sn * sn, 0.0);
    el.fml.assign(nf * sn * nk, 0.0);
    for (int f = 0; f < nf; ++f)
      for (int g = 0; g < sg; ++g) {
        const double* ph = &el.fphi[(f * sg + g) * sn];
        for (int i = 0; i < sn; ++i) {
          for (int j = 0; j < sn; ++j) el.fmf[(f * sn + i) * sn + j] += fw[g] * ph[i] * ph[j];
          for (int k = 0; k < nk; ++k)
            el.fml[(f * sn + i) * nk + k] += fw[g] * ph[i] * el.flk[g * nk + k];
        }
      }
  }
  // volume term through the quadratic gradient (P2 nodes: vertices, then edge midpoints)
  {
    std::vector<std::array<int, 2>> p2;
    for (int i = 0; i < nb; ++i) p2.push_back({i, i});
    for (int i = 0; i < nb; ++i)
      for (int j = i + 1; j < nb; ++j) p2.push_back({i, j});
    const int np2 = (int)p2.size();
    el.np2 = np2;
    auto psi = [&](const double* lam, double* out) {
      for (int i = 0; i < np2; ++i) {
        const int a = p2[i][0], b = p2[i][1];
        out[i] = a == b ? lam[a] * (2.0 * lam[a] - 1.0) : 4.0 * lam[a] * lam[b];
      }
    };
    el.vd.assign(D * np2 * nloc, 0.0);
    for (int i = 0; i < np2; ++i) {
      double lam[4] = {0, 0, 0, 0};
      lam[p2[i][0]] += 0.5;
      lam[p2[i][1]] += 0.5;
      lagrange(el.index, nb, lam, phi.data(), dxi.data(), nloc);
      for (int d = 0; d < D; ++d)
        for (int n = 0; n < nloc; ++n) el.vd[(d * np2 + i) * nloc + n] = dxi[d * nloc + n];
    }
    el.vm2.assign(np2 * np2, 0.0);
    std::vector<double> ps(np2);
    for (int g = 0; g < el.ngi; ++g) {
      psi(&vlam[g * nb], ps.data());
      for (int i = 0; i < np2; ++i)
        for (int j = 0; j < np2; ++j) el.vm2[i * np2 + j] += vw[g] * ps[i] * ps[j];
    }
    double err = 0.0, big = 0.0;
    std::vector<double> t(np2);
    for (int d = 0; d < D; ++d)
      for (int e2 = 0; e2 < D; ++e2)
        for (int n = 0; n < nloc; ++n)
          for (int m = 0; m < nloc; ++m) {
            double ref = 0.0;
            for (int g = 0; g < el.ngi; ++g)
              ref += vw[g] * el.nlx[(d * nloc + n) * el.ngi + g] * el.nlx[(e2 * nloc + m) * el.ngi + g];
            double v = 0.0;
            for (int i = 0; i < np2; ++i)
              for (int j = 0; j < np2; ++j)
                v += el.vd[(d * np2 + i) * nloc + n] * el.vm2[i * np2 + j] *
                     el.vd[(e2 * np2 + j) * nloc + m];
            err = std::max(err, std::fabs(v - ref));
            big = std::max(big, std::fabs(ref));
          }
    el.vol_factor_err = big > 0.0 ? err / big : err;
  }
  el.p1_prolong.assign(nloc * nb, 0.0);
  for (int n = 0; n < nloc; ++n)
    for (int k = 0; k < nb; ++k) el.p1_prolong[n * nb + k] = el.index[n * nb + k] / double(P);
  if (D == 3) el.drst.assign(&TET_DRST[0][0][0], &TET_DRST[0][0][0] + 24);
  else el.drst.assign(&TRI_DRST[0][0], &TRI_DRST[0][0] + 6);
  if (D == 3) el.face_verts.assign(&TET_FACE_VERTS[0][0], &TET_FACE_VERTS[0][0] + 12);
  else el.face_verts.assign(&TRI_FACE_VERTS[0][0], &TRI_FACE_VERTS[0][0] + 6);
  return el;
}

// elements of every vertex, in increasing element order (CSR)
static void vertex_elements(const int64_t* cells, int64_t nele, int nb, int64_t nverts,
                            std::vector<int64_t>& vptr, std::vector<int64_t>& velem) {
  vptr.assign(nverts + 1, 0);
  for (int64_t q = 0; q < nele * nb; ++q) vptr[cells[q] + 1]++;
  for (int64_t v = 0; v < nverts; ++v) vptr[v + 1] += vptr[v];
  velem.resize(nele * nb);
  std::vector<int64_t> fill(vptr.begin(), vptr.end() - 1);
  for (int64_t e = 0; e < nele; ++e)
    for (int k = 0; k < nb; ++k) velem[fill[cells[e * nb + k]]++] = e;
}

Topology build_topology(const int64_t* cells, int64_t nele, int D) {
  // Two faces match when their vertex sets agree.  Faces are bucketed by their smallest
  // vertex (a counting sort), each bucket is sorted by the remaining vertices, and equal
  // neighbours in a bucket are paired.  The alignment of a matched pair is read from the
  // pair (lower global face id first) and stored on both sides.
  const int nf = D + 1, nb = D + 1;
  const int64_t nfaces = nele * nf;
  Topology t;
  t.nbf.assign(nfaces, -1);
  t.alnmt.assign(nfaces, -1);
  t.bctype.assign(nfaces, 0);
  int64_t nverts = 0;
  for (int64_t q = 0; q < nele * nb; ++q) nverts = std::max(nverts, cells[q] + 1);
  auto face_verts = [&](int64_t z, int64_t* v) {
    const int64_t e = z / nf;
    const int f = (int)(z - e * nf);
    for (int i = 0; i < D; ++i)
      v[i] = cells[e * nb + (D == 3 ? TET_FACE_VERTS[f][i] : TRI_FACE_VERTS[f][i])];
  };
  // faces bucketed by their smallest vertex, each record holding the other two sorted
  // vertices (D == 2: the second is 0) and the face id
  struct Rec {
    int64_t k1, k2, z;
  };
  std::vector<int64_t> bptr(nverts + 1, 0);
  std::vector<int32_t> kmin(nfaces);
  parallel_chunks(nfaces, [&](int, int64_t a, int64_t b) {
    for (int64_t z = a; z < b; ++z) {
      int64_t v[3] = {0, 0, 0};
      face_verts(z, v);
      kmin[z] = (int32_t)*std::min_element(v, v + D);
    }
  });
  for (int64_t z = 0; z < nfaces; ++z) bptr[kmin[z] + 1]++;
  for (int64_t v = 0; v < nverts; ++v) bptr[v + 1] += bptr[v];
  std::vector<Rec> rec(nfaces);
  {
    std::vector<int64_t> fill(bptr.begin(), bptr.end() - 1);
    for (int64_t z = 0; z < nfaces; ++z) rec[fill[kmin[z]]++].z = z;
  }
  parallel_chunks(nverts, [&](int, int64_t va, int64_t vb) {
    for (int64_t v = va; v < vb; ++v) {
      Rec* lo = rec.data() + bptr[v];
      Rec* hi = rec.data() + bptr[v + 1];
      for (Rec* r = lo; r < hi; ++r) {
        int64_t w[3] = {0, 0, 0};
        face_verts(r->z, w);
        std::sort(w, w + D);
        r->k1 = w[1];
        r->k2 = D == 3 ? w[2] : 0;
      }
      std::sort(lo, hi, [](const Rec& x, const Rec& y) {
        return x.k1 != y.k1 ? x.k1 < y.k1 : (x.k2 != y.k2 ? x.k2 < y.k2 : x.z < y.z);
      });
      for (Rec* p = lo; p + 1 < hi; ++p) {
        if (p[0].k1 != p[1].k1 || p[0].k2 != p[1].k2) continue;
        const int64_t z = p[0].z, z2 = p[1].z;
        int64_t va3[3] = {0, 0, 0}, vb3[3] = {0, 0, 0};
        face_verts(z, va3);  // z < z2: the lower global face id first
        face_verts(z2, vb3);
        const int64_t al = va3[0] == vb3[0] ? 0 : (va3[1] == vb3[0] ? 1 : 2);
        t.nbf[z] = z2;
        t.nbf[z2] = z;
        t.alnmt[z] = t.alnmt[z2] = al;
        ++p;  // a face pairs once
      }
    }
  });
  for (int64_t z = 0; z < nfaces; ++z) t.bctype[z] = t.nbf[z] >= 0 ? -1 : 0;
  return t;
}

void p1cg_sparsity(const int64_t* cells, int64_t nele, int D, int64_t nverts,
                   std::vector<int64_t>& fina, std::vector<int64_t>& cola) {
  // row i = sorted distinct vertices of the elements around vertex i
  const int nb = D + 1;
  std::vector<int64_t> vptr, velem;
  vertex_elements(cells, nele, nb, nverts, vptr, velem);
  auto row = [&](int64_t v, std::vector<int64_t>& buf) {
    buf.clear();
    for (int64_t q = vptr[v]; q < vptr[v + 1]; ++q)
      for (int k = 0; k < nb; ++k) buf.push_back(cells[velem[q] * nb + k]);
    std::sort(buf.begin(), buf.end());
    buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
  };
  fina.assign(nverts + 1, 0);
  parallel_chunks(nverts, [&](int, int64_t a, int64_t b) {
    std::vector<int64_t> buf;
    for (int64_t v = a; v < b; ++v) {
      row(v, buf);
      fina[v + 1] = (int64_t)buf.size();
    }
  });
  for (int64_t v = 0; v < nverts; ++v) fina[v + 1] += fina[v];
  cola.resize(fina[nverts]);
  parallel_chunks(nverts, [&](int, int64_t a, int64_t b) {
    std::vector<int64_t> buf;
    for (int64_t v = a; v < b; ++v) {
      row(v, buf);
      std::copy(buf.begin(), buf.end(), cola.begin() + fina[v]);
    }
  });
}

int color2_range(const std::vector<int64_t>& fina, const std::vector<int64_t>& cola, int64_t v0,
                 int64_t v1, std::vector<int32_t>& whichc) {
  // First-fit in vertex order: vertex i takes the smallest colour that no distance-1 or
  // distance-2 neighbour before it holds.  This is the colouring of the reference's
  // multi-pass loop (z02/sparsity.py:235-269), which gives colour c in pass c to every
  // uncoloured vertex that no earlier vertex of colour c is near: both decide vertex i from
  // the vertices before it only.  Rows of [v0, v1) must reference columns in [v0, v1).
  std::vector<int64_t> seen(64, -1);  // seen[c] == i: colour c is taken near vertex i
  int ncolor = 0;
  for (int64_t i = v0; i < v1; ++i) {
    for (int64_t c = fina[i]; c < fina[i + 1]; ++c) {
      const int64_t j = cola[c];
      const int32_t cj = whichc[j];
      if (cj > 0) seen[cj] = i;
      for (int64_t c2 = fina[j]; c2 < fina[j + 1]; ++c2) {
        const int32_t ck = whichc[cola[c2]];
        if (ck > 0) seen[ck] = i;
      }
    }
    int32_t k = 1;
    while (k < (int32_t)seen.size() && seen[k] == i) ++k;
    if (k >= (int32_t)seen.size()) seen.resize(2 * seen.size(), -1);
    whichc[i] = k;
    ncolor = std::max(ncolor, (int)k);
  }
  return ncolor;
}

int color2(const std::vector<int64_t>& fina, const std::vector<int64_t>& cola,
           std::vector<int32_t>& whichc) {
  const int64_t n = (int64_t)fina.size() - 1;
  whichc.assign(n, 0);
  return n > 0 ? color2_range(fina, cola, 0, n, whichc) : 0;
}

void hilbert_order(const double* verts, int64_t nv, int D, std::vector<int64_t>& sfc) {
  const int bits = D == 3 ? 10 : 16;
  double lo[3], hi[3];
  for (int d = 0; d < D; ++d) {
    lo[d] = 1e300;
    hi[d] = -1e300;
  }
  for (int64_t i = 0; i < nv; ++i)
    for (int d = 0; d < D; ++d) {
      lo[d] = std::min(lo[d], verts[i * D + d]);
      hi[d] = std::max(hi[d], verts[i * D + d]);
    }
  const double scale = double((1LL << bits) - 1);
  std::vector<int64_t> key(nv);
  for (int64_t i = 0; i < nv; ++i) {
    int64_t X[3] = {0, 0, 0};
    for (int d = 0; d < D; ++d) {
      double span = hi[d] > lo[d] ? hi[d] - lo[d] : 1.0;
      double t = (verts[i * D + d] - lo[d]) / span;
      t = t * scale;
      t = t + 0.5;
      X[d] = (int64_t)std::floor(t);
    }
    // Skilling, "Programming the Hilbert curve" (AIP Conf. Proc. 707, 2004)
    const int64_t M = 1LL << (bits - 1);
    for (int64_t Q = M; Q > 1; Q >>= 1) {
      int64_t Pm = Q - 1;
      for (int k = 0; k < D; ++k) {
        if (X[k] & Q) {
          X[0] ^= Pm;
        } else {
          int64_t t = (X[0] ^ X[k]) & Pm;
          X[0] ^= t;
          X[k] ^= t;
        }
      }
    }
    for (int k = 1; k < D; ++k) X[k] ^= X[k - 1];
    int64_t t = 0;
    for (int64_t Q = M; Q > 1; Q >>= 1)
      if (X[D - 1] & Q) t ^= Q - 1;
    for (int k = 0; k < D; ++k) X[k] ^= t;
    int64_t kk = 0;
    for (int j = bits - 1; j >= 0; --j)
      for (int k = 0; k < D; ++k) kk = (kk << 1) | ((X[k] >> j) & 1);
    key[i] = kk;
  }
  std::vector<int64_t> order(nv);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int64_t a, int64_t b) { return key[a] < key[b]; });
  sfc.assign(nv, 0);
  for (int64_t p = 0; p < nv; ++p) sfc[order[p]] = p + 1;
}

int nlevel_sfc(int64_t n) {
  if (n == 1) return 1;
  int lev = 1;
  while (n > 1) {
    n = (n - 1) / 2 + 1;
    ++lev;
  }
  return lev;
}

void build_sfc_level_pattern(const std::vector<int64_t>& fina, const std::vector<int64_t>& cola,
                             const std::vector<int64_t>& sfc, SfcLevelPattern& pat) {
  // Level 0 is the pattern renumbered into SFC order, each row sorted by column.  Row I of
  // level l+1 gathers rows 2I and 2I+1 of level l with columns c -> c >> 1 (positions p
  // and q of one level are summed into p >> 1, q >> 1: the aggregation of
  // compact_fcode.f90:7-225); entries with one column are summed in the order: row 2I by
  // column, then row 2I+1 by column.  Rows are independent: each pass counts, then fills,
  // in parallel.  Only the sources of each entry are recorded here.
  const int64_t n0 = (int64_t)fina.size() - 1;
  pat.nlevel = nlevel_sfc(n0);
  pat.n.assign(pat.nlevel, 0);
  pat.rowptr.assign(pat.nlevel, {});
  pat.col.assign(pat.nlevel, {});
  pat.sptr.assign(pat.nlevel, {});
  pat.src.assign(pat.nlevel, {});
  // ---- level 0
  {
    std::vector<int64_t> node(n0);  // node at each SFC position
    for (int64_t i = 0; i < n0; ++i) node[sfc[i] - 1] = i;
    std::vector<int32_t>& rp = pat.rowptr[0];
    rp.assign(n0 + 1, 0);
    for (int64_t p = 0; p < n0; ++p) rp[p + 1] = (int32_t)(fina[node[p] + 1] - fina[node[p]]);
    for (int64_t p = 0; p < n0; ++p) rp[p + 1] += rp[p];
    pat.col[0].resize(rp[n0]);
    pat.perm0.resize(rp[n0]);
    parallel_chunks(n0, [&](int, int64_t a, int64_t b) {
      std::vector<std::pair<int32_t, int32_t>> row;
      for (int64_t p = a; p < b; ++p) {
        const int64_t i = node[p];
        row.clear();
        for (int64_t c = fina[i]; c < fina[i + 1]; ++c)
          row.push_back({(int32_t)(sfc[cola[c]] - 1), (int32_t)c});
        std::stable_sort(row.begin(), row.end(),
                         [](const auto& x, const auto& y) { return x.first < y.first; });
        for (size_t k = 0; k < row.size(); ++k) {
          pat.col[0][rp[p] + k] = row[k].first;
          pat.perm0[rp[p] + k] = row[k].second;
        }
      }
    });
    pat.n[0] = n0;
  }
  // ---- coarser levels
  for (int l = 1; l < pat.nlevel; ++l) {
    const int64_t nf = pat.n[l - 1], nc = (nf - 1) / 2 + 1;
    const std::vector<int32_t>& frp = pat.rowptr[l - 1];
    const std::vector<int32_t>& fcol = pat.col[l - 1];
    std::vector<int32_t>& rp = pat.rowptr[l];
    rp.assign(nc + 1, 0);
    // the coarse row of the pair as (column, source entry), sorted by column (stable)
    auto merge = [&](int64_t I, std::vector<std::pair<int32_t, int32_t>>& row) {
      row.clear();
      for (int64_t i = 2 * I; i < std::min(2 * I + 2, nf); ++i)
        for (int32_t k = frp[i]; k < frp[i + 1]; ++k) row.push_back({fcol[k] >> 1, k});
      std::stable_sort(row.begin(), row.end(),
                       [](const auto& x, const auto& y) { return x.first < y.first; });
    };
    auto distinct = [](const std::vector<std::pair<int32_t, int32_t>>& row) {
      int32_t m = 0;
      for (size_t k = 0; k < row.size(); ++k)
        if (k == 0 || row[k].first != row[k - 1].first) ++m;
      return m;
    };
    std::vector<int32_t> scount(nc + 1, 0);  // source entries per coarse row
    parallel_chunks(nc, [&](int, int64_t a, int64_t b) {
      std::vector<std::pair<int32_t, int32_t>> row;
      for (int64_t I = a; I < b; ++I) {
        merge(I, row);
        rp[I + 1] = distinct(row);
        scount[I + 1] = (int32_t)row.size();
      }
    });
    for (int64_t I = 0; I < nc; ++I) {
      rp[I + 1] += rp[I];
      scount[I + 1] += scount[I];
    }
    pat.col[l].resize(rp[nc]);
    pat.sptr[l].resize((size_t)rp[nc] + 1);
    pat.src[l].resize(scount[nc]);
    pat.sptr[l][rp[nc]] = scount[nc];
    parallel_chunks(nc, [&](int, int64_t a, int64_t b) {
      std::vector<std::pair<int32_t, int32_t>> row;
      for (int64_t I = a; I < b; ++I) {
        merge(I, row);
        int32_t w = rp[I] - 1, s = scount[I];
        for (size_t k = 0; k < row.size(); ++k) {
          if (k == 0 || row[k].first != row[k - 1].first) {
            ++w;
            pat.col[l][w] = row[k].first;
            pat.sptr[l][w] = s;
          }
          pat.src[l][s++] = row[k].second;
        }
      }
    });
    pat.n[l] = nc;
  }
}

void fill_sfc_levels(SfcLevelPattern& pat, const std::vector<double>& rar, SfcLevelsHost& lv) {
  lv.nlevel = pat.nlevel;
  lv.n = pat.n;
  lv.rowptr = std::move(pat.rowptr);
  lv.col = std::move(pat.col);
  lv.val.assign(lv.nlevel, {});
  lv.diag.assign(lv.nlevel, {});
  for (int l = 0; l < lv.nlevel; ++l) {
    const int64_t n = lv.n[l];
    const std::vector<int32_t>& rp = lv.rowptr[l];
    std::vector<double>& v = lv.val[l];
    v.resize(rp[n]);
    if (l == 0) {
      parallel_chunks((int64_t)v.size(), [&](int, int64_t a, int64_t b) {
        for (int64_t k = a; k < b; ++k) v[k] = rar[pat.perm0[k]];
      }, 1 << 16);
    } else {
      const std::vector<double>& fv = lv.val[l - 1];
      const std::vector<int32_t>& sp = pat.sptr[l];
      const std::vector<int32_t>& sr = pat.src[l];
      parallel_chunks((int64_t)v.size(), [&](int, int64_t a, int64_t b) {
        for (int64_t k = a; k < b; ++k) {
          double s = fv[sr[sp[k]]];
          for (int32_t j = sp[k] + 1; j < sp[k + 1]; ++j) s += fv[sr[j]];
          v[k] = s;
        }
      }, 1 << 16);
    }
    lv.diag[l].assign(n, 0.0);
    parallel_chunks(n, [&](int, int64_t a, int64_t b) {
      for (int64_t i = a; i < b; ++i)
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k)
          if (lv.col[l][k] == i) lv.diag[l][i] = v[k];
    });
  }
}

void build_sfc_levels(const std::vector<int64_t>& fina, const std::vector<int64_t>& cola,
                      const std::vector<double>& rar, const std::vector<int64_t>& sfc,
                      SfcLevelsHost& lv) {
  SfcLevelPattern pat;
  build_sfc_level_pattern(fina, cola, sfc, pat);
  fill_sfc_levels(pat, rar, lv);
}

}  // namespace dgmi
